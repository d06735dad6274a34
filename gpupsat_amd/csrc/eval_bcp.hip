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

// Orders this wavefront's accesses to its mask table: LDS (wave_sync), or, when the table
// sits in an HBM slice (instances too large for LDS), after waiting for every memory
// operation of the wavefront to complete.
template <bool GTAB>
__device__ __forceinline__ void table_sync()
{
    if (GTAB) __builtin_amdgcn_s_waitcnt(0);
    wave_sync();
}

// Stage the instance's A rows (A*W words) and build LT (2 words per literal slot).
template <bool GTAB>
__device__ __forceinline__ void build_literal_masks(const uint32_t* __restrict__ rows, int A, int W,
                                                    uint32_t* stage, uint32_t* lt32, int lane, const Transposer& tp)
{
    bulk_to_lds(stage, rows, A * W, lane);
    table_sync<GTAB>();
    const int half = lane >> 5, i = lane & 31;
    for (int j = 0; j < W; j++) {
        uint32_t x = (lane < A) ? stage[lane * W + j] : 0u;
        x = tp(x);
        lt32[2 * (32 * j + i) + half] = x;
    }
    table_sync<GTAB>();
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
// (low word) and 32+r (high word), the statuses of the 32 clauses of its half.  The
// accumulators count b0, b1 and UNIT = b0 & b1; finish_counts turns them into
// #SAT = |b0| - #UNIT and #UNSAT = |b1| - #UNIT.
__device__ __forceinline__ void count_group(uint64_t b0, uint64_t b1, const Transposer& tp, uint32_t (&acc)[2][3])
{
    const uint32_t w0[2] = {tp((uint32_t)b0), tp((uint32_t)(b0 >> 32))};
    const uint32_t w1[2] = {tp((uint32_t)b1), tp((uint32_t)(b1 >> 32))};
#pragma unroll
    for (int h = 0; h < 2; h++) {
        acc[h][0] += __builtin_popcount(w0[h]);
        acc[h][1] += __builtin_popcount(w1[h]);
        acc[h][2] += __builtin_popcount(w0[h] & w1[h]);
    }
}

// Per-assignment #SAT, #UNSAT, #UNIT from count_group's accumulators: assignment a sums
// the low words of lanes a and a^32 (a < 32) or the high words (a >= 32).
__device__ __forceinline__ void finish_counts(const uint32_t (&acc)[2][3], int lane, uint32_t (&tot)[3])
{
#pragma unroll
    for (int s = 0; s < 3; s++) {
        const uint32_t lo = acc[0][s] + shfl_xor32(acc[0][s]);
        const uint32_t hi = acc[1][s] + shfl_xor32(acc[1][s]);
        tot[s] = (lane < 32) ? lo : hi;
    }
    tot[0] -= tot[2];
    tot[1] -= tot[2];
}

// ---------------------------------------------------------------------------------------
// DO_EVAL: write #SAT/#UNSAT/#UNIT per (instance, assignment) for the input rows.
// DO_BCP:  write the closure rows, conflict flags and implied-literal counts.
// STAGE: the instance's clause arena (literals and offsets) is copied to LDS first, so
// every clause group reads LDS only.
// GTAB: the mask table and row staging sit in the instance's slice of the HBM buffer
// gtab (lds_words words per instance) instead of LDS: instances whose 2n literal masks do
// not fit a workgroup's LDS.
template <bool DO_EVAL, bool DO_BCP, bool STAGE, bool GTAB>
__global__ __launch_bounds__(256) void k_eval_bcp(const InstMeta* __restrict__ meta, const int32_t* __restrict__ lits,
                                                   const int32_t* __restrict__ coff, const uint32_t* __restrict__ rows,
                                                   int A, int32_t* __restrict__ counts, uint32_t* __restrict__ out_rows,
                                                   int32_t* __restrict__ conflict, int32_t* __restrict__ props,
                                                   int32_t* __restrict__ passes_out, int n_inst, int lds_words,
                                                   uint32_t* __restrict__ gtab)
{
    extern __shared__ __align__(16) uint32_t lds_all[];
    // the wavefront's slot is uniform: its metadata then comes in with scalar loads
    const int wslot = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int inst = blockIdx.x * (blockDim.x / kWave) + wslot;
    const int lane = threadIdx.x % kWave;
    if (inst >= n_inst) return;
    uint32_t* smem = GTAB ? gtab + (size_t)inst * lds_words : lds_all + (size_t)wslot * lds_words;
    const Transposer tp(lane);
    const InstMeta mi = meta[inst];
    const int W = __builtin_amdgcn_readfirstlane(words_per_row(mi.n));
    const int m = __builtin_amdgcn_readfirstlane(mi.m);
    uint32_t* lt32 = smem;                       // 64*W words (2 per literal slot)
    uint32_t* stage = smem + 64 * W;             // A*W words
    uint32_t* conf = stage + 64 * W;             // 4 spare words before the staged arena
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

    build_literal_masks<GTAB>(rows + row0, A, W, stage, lt32, lane, tp);
    table_sync<GTAB>();

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
        uint32_t tot[3];
        finish_counts(acc, lane, tot);
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
        uint64_t frozen;
        {
            uint64_t cm = 0;
            for (int v = lane; v < 16 * W; v += kWave) {
                const uint4 q = lt4[v];
                cm |= ((uint64_t)q.x | ((uint64_t)q.y << 32)) & ((uint64_t)q.z | ((uint64_t)q.w << 32));
            }
            frozen = wave_or64(cm);
        }
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
                    frozen |= wave_or64(newconf);
                    if ((frozen & amask) == amask) break;
                }
            }
            changed = ballot(my_change) != 0ull && (frozen & amask) != amask;
            table_sync<GTAB>();
        }
        // complementary implications are conflicts too
        uint64_t cm = 0;
        for (int v = lane; v < 16 * W; v += kWave) {
            const uint4 q = lt4[v];
            cm |= ((uint64_t)q.x | ((uint64_t)q.y << 32)) & ((uint64_t)q.z | ((uint64_t)q.w << 32));
        }
        const uint64_t cmask = frozen | wave_or64(cm);
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

// ---------------------------------------------------------------------------------------
// Register-resident path for instances whose clauses all hold exactly three literals,
// with n <= 256 (W <= 16 words per row) and m <= 1024 (<= 16 clause groups).  The clauses
// come from the batch's packed clause words (three 10-bit literal codes per 4-byte word).
// Every global load of the instance (its A rows and its m clause words) is issued before
// the first one is used; each lane then keeps its 16 clause words in VGPRs, so the BCP
// passes touch LDS only.  With BCP, the input rows are copied to the output as they
// arrive; only rows that gain implications are written again.
//
// LDS holds the literal masks twice: slot l (16 bytes) = {T(l), T(l ^ 1)}, the true and
// the false mask of literal l, so a clause literal costs one 16-byte read and no select.
// A propagated literal l is OR-ed into T(l) of slot l and into the false half of slot
// l ^ 1.  The first 64*W words double as the row staging area; slot 32*W is all zero.
// With at most 32 open rows (from the start, or once the 64-bit passes are down to 8-32
// rows still gaining implications), BCP runs on a compacted table of 8-byte slots
// {T(l), T(l ^ 1)} over those rows, at words [64*W, 128*W + 2).
// LDS: (128*W + 8) words, so every wavefront's region starts 32-byte aligned (the BCP passes
// find a slot's partner l ^ 1 by flipping one address bit).
// Byte address of literal slot FIELD (0-2) of a packed clause word p in a table of
// (1 << SHIFT)-byte slots at base: one bit-field extract and one shift-add (written out,
// since the compiler otherwise turns the pair into a shift, a mask and an add).  The base is
// wave-uniform and read from its scalar register (a vector operand cost one v_mov per use).
template <int FIELD, int SHIFT>
__device__ __forceinline__ uint32_t slot_addr(uint32_t base, uint32_t p)
{
    uint32_t t, r;
    asm("v_bfe_u32 %0, %1, %2, 10" : "=v"(t) : "v"(p), "i"(10 * FIELD));
    asm("v_lshl_add_u32 %0, %1, %2, %3" : "=v"(r) : "v"(t), "i"(SHIFT), "s"(base));
    return r;
}

__device__ __forceinline__ int64_t uni64(int64_t x)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

constexpr int kFastW = kFast3MaxWords;
constexpr int kFastG = kFast3MaxGroups;

__device__ __forceinline__ uint64_t u64(uint32_t lo, uint32_t hi) { return (uint64_t)lo | ((uint64_t)hi << 32); }

// Symmetric 3-input bit functions as one v_bitop3_b32 each (the truth tables are invariant
// under any order of the three sources): at least two of a, b, c set; exactly two set;
// all three set.
__device__ __forceinline__ uint32_t maj3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t d;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xe8" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}
__device__ __forceinline__ uint32_t two3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t d;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x68" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}
__device__ __forceinline__ uint32_t and3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t d;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x80" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}
// a & ~b & ~c in one instruction (the compiler otherwise forms ~(b | c) and then the AND, and
// keeps the complement live for the second use)
__device__ __forceinline__ uint32_t andn2(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t d;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x10" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// a native 4-word vector: 16-byte LDS reads of it stay single ds_read_b128 instructions
// (the HIP uint4 struct is split into 8-byte halves by the compiler when its members are
// consumed as 64-bit pairs)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// LDS accesses by byte address: a BCP pass forms each clause literal's slot address once and
// uses it for the slot read and for both implication writes.  The partner slot l ^ 1 is the
// address with bit 3 flipped in a table of 8-byte slots starting 16-byte aligned, and with
// bit 4 flipped in a table of 16-byte slots starting 32-byte aligned.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) u32x2 lds_u32x2;
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;
__device__ __forceinline__ uint32_t lds_addr(const void* p)
{
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
__device__ __forceinline__ uint2 lds_read8(uint32_t a)
{
    const u32x2 v = *(const lds_u32x2*)(uintptr_t)a;
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ u32x4 lds_read16(uint32_t a) { return *(const lds_u32x4*)(uintptr_t)a; }
__device__ __forceinline__ void lds_or32(uint32_t a, uint32_t v)
{
    __hip_atomic_fetch_or((lds_u32*)(uintptr_t)a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_or64(uint32_t a, uint64_t v)
{
    __hip_atomic_fetch_or((lds_u64*)(uintptr_t)a, (unsigned long long)v, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <bool DO_EVAL, bool DO_BCP>
__global__ __launch_bounds__(256) void k_eval_bcp3(const InstMeta* __restrict__ meta, const uint32_t* __restrict__ pk3,
                                                    const uint32_t* __restrict__ rows, int A, int32_t* __restrict__ counts,
                                                    uint32_t* __restrict__ out_rows, int32_t* __restrict__ conflict,
                                                    int32_t* __restrict__ props, int32_t* __restrict__ passes_out,
                                                    int n_inst, int lds_words, int small64)
{
    extern __shared__ __align__(32) uint32_t lds_all[];
    // the wavefront's slot is uniform: its metadata then comes in with scalar loads
    const int wslot = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int inst = blockIdx.x * (blockDim.x / kWave) + wslot;
    const int lane = threadIdx.x % kWave;
    if (inst >= n_inst) return;
    uint32_t* smem = lds_all + (size_t)wslot * lds_words;
    const InstMeta mi = meta[inst];
    // the instance's metadata in SGPRs: the group loops branch on scalars, and the
    // addresses of all loads are known before any of them is waited for
    const int W = __builtin_amdgcn_readfirstlane(words_per_row(mi.n));
    const int m = __builtin_amdgcn_readfirstlane(mi.m);
    const int64_t clause_base = uni64(mi.coff_base - inst);
    const int64_t word_base = uni64(mi.word_base);
    const int G = (m + kWave - 1) / kWave;
    // slot 32*W stays zero: lanes past m evaluate a clause of three such literals, which
    // has no status and no unit, so no group needs a validity mask
    const uint32_t zslot = 32u * (uint32_t)W;
    const uint32_t zpk = zslot | (zslot << 10) | (zslot << 20);
    uint32_t* dup = smem;                        // 32*W slots of 4 words, then the zero slot
    const u32x4* dup4 = reinterpret_cast<const u32x4*>(dup);
    const uint32_t dupb = lds_addr(dup);         // its LDS byte address (32-byte aligned)
    const int64_t row0 = (int64_t)A * word_base;
    const uint32_t* ipk = pk3 + clause_base;

    // rows: coalesced copy of the A*W words into LDS; clause words: straight into VGPRs.
    // Every load sits in a scalar-guarded block with no use inside it (lanes past the end
    // re-read the last word), so all of them are in flight together.
    const uint32_t* src = rows + row0;
    const int nw = A * W;
    uint32_t rbuf[kFastW];
#pragma unroll
    for (int u = 0; u < kFastW; u++) {
        rbuf[u] = 0u;
        if (u * kWave < nw) rbuf[u] = __builtin_nontemporal_load(src + u * kWave + min(lane, nw - 1 - u * kWave));
    }
    // clause words: one scalar-guarded block per group with no use inside it, so all G
    // loads are in flight together (lanes past m re-read clause m-1, which stays in bounds
    // and is replaced by the zero triple below)
    uint32_t pk[kFastG];
    const int mc = m > 0 ? m - 1 : 0;
#pragma unroll
    for (int g = 0; g < kFastG; g++) {
        pk[g] = 0u;
        if (g < G) pk[g] = __builtin_nontemporal_load(ipk + g * kWave + min(lane, mc - g * kWave));
    }
    // a conflicting row's output is its input: every output row starts as a copy (the
    // rows that gain implications are rewritten at the end).  Word offsets are split
    // into a scalar part and the lane, so no per-word lane index stays live.  The copy and
    // the rewrite are nontemporal stores: they leave L2 to the launch's opening burst of
    // loads (p = 0.3: 59 -> 56.7 us; at p = 0.02-0.1 unchanged to -1%).
    uint32_t* dst = out_rows + row0;
#pragma unroll
    for (int u = 0; u < kFastW; u++) {
        if (u * kWave < nw && lane < nw - u * kWave) {
            dup[u * kWave + lane] = rbuf[u];
            if (DO_BCP) __builtin_nontemporal_store(rbuf[u], dst + u * kWave + lane);
        }
    }
    wave_sync();
    // the rows are needed only to build the literal table; afterwards each lane keeps
    // its row's literal count
    uint32_t in[kFastW];
    int in_pop = 0;
#pragma unroll
    for (int j = 0; j < kFastW; j++) {
        in[j] = (j < W && lane < A) ? dup[lane * W + j] : 0u;
        in_pop += __builtin_popcount(in[j]);
    }
    // the count is formed here: sunk to its use at the end, it would keep all W words live
    asm volatile("" : "+v"(in_pop));
    wave_sync();
    const Transposer tp(lane);
    const int half = lane >> 5, i = lane & 31;
    if (lane < 4) dup[4 * zslot + lane] = 0u;
#pragma unroll
    for (int j = 0; j < kFastW; j++) {
        if (j < W) {
            const uint32_t x = tp(in[j]);
            const int l = 32 * j + i;
            dup[4 * l + half] = x;
            dup[4 * (l ^ 1) + 2 + half] = x;
        }
    }
    // lanes past m take the zero triple; the empty asm keeps the select out of the groups
#pragma unroll
    for (int g = 0; g < kFastG; g++) {
        pk[g] = (lane < m - g * kWave) ? pk[g] : zpk;
        asm volatile("" : "+v"(pk[g]));
    }
    wave_sync();

    uint64_t unsat_any = 0;      // rows with a clause falsified by the input
    if (DO_EVAL) {
        uint32_t acc[2][3] = {{0, 0, 0}, {0, 0, 0}};
        // software pipeline: the three 16-byte reads of group g+1 are issued before group
        // g is evaluated.  Two named buffers alternate, so the unrolled groups need no
        // register copies where a group's scalar guard joins.
        const auto fetch = [&](uint32_t p, u32x4& a, u32x4& b, u32x4& c) {
            a = lds_read16(slot_addr<0, 4>(dupb, p));
            b = lds_read16(slot_addr<1, 4>(dupb, p));
            c = lds_read16(slot_addr<2, 4>(dupb, p));
        };
        // per clause: SAT|UNIT = T | (exactly two literals false), UNSAT|UNIT = (at least
        // two false) & ~T, two operations per 32-bit word each
        const auto eval_group = [&](const u32x4& qa, const u32x4& qb, const u32x4& qc) {
            const uint32_t Tl = qa.x | qb.x | qc.x, Th = qa.y | qb.y | qc.y;
            const uint32_t b0l = Tl | two3(qa.z, qb.z, qc.z), b0h = Th | two3(qa.w, qb.w, qc.w);
            const uint32_t b1l = maj3(qa.z, qb.z, qc.z) & ~Tl, b1h = maj3(qa.w, qb.w, qc.w) & ~Th;
            count_group(u64(b0l, b0h), u64(b1l, b1h), tp, acc);
        };
        u32x4 xa, xb, xc, ya, yb, yc;
        fetch(pk[0], xa, xb, xc);
#pragma unroll
        for (int g = 0; g < kFastG; g += 2) {
            if (g < G) {
                if (g + 1 < kFastG) fetch(pk[g + 1], ya, yb, yc);
                eval_group(xa, xb, xc);
            }
            if (g + 1 < G) {
                if (g + 2 < kFastG) fetch(pk[g + 2], xa, xb, xc);
                eval_group(ya, yb, yc);
            }
        }
        uint32_t tot[3];
        finish_counts(acc, lane, tot);
        if (lane < A) {
            int32_t* o = counts + ((int64_t)inst * A + lane) * 3;
            o[0] = (int32_t)tot[0];
            o[1] = (int32_t)tot[1];
            o[2] = (int32_t)tot[2];
        }
        unsat_any = ballot(lane < A && tot[1] != 0u);
    }

    if (DO_BCP) {
        const uint64_t amask = A >= 64 ? ~0ull : ((1ull << A) - 1ull);
        // contradictory inputs, and inputs that already falsify a clause, are conflicts
        uint64_t frozen0;
        {
            uint64_t cm = unsat_any;
            for (int v = lane; v < 16 * W; v += kWave) {
                const u32x4 q = dup4[2 * v];
                cm |= u64(q.x, q.y) & u64(q.z, q.w);
            }
            frozen0 = wave_or64(cm);
        }
        uint64_t frozen = frozen0;
        int pass = 0;
        bool my_conflict = (frozen >> lane) & 1ull;
        int cnt = 0;            // literals in the closure of an open row
        uint64_t act = amask & ~frozen;
        int n_act = __builtin_popcountll(act);
        // More than 32 open rows: passes over the 64-bit table.  A row that gains no
        // implication during a whole pass is at its fixpoint (every clause was evaluated
        // against its final state).  When, from the second pass on, 8 to 32 rows still gain
        // implications, the others' closures are staged for output and the passes go on over a
        // compacted table of the rows still open (a handoff).
        bool handoff = false;
        int a_k = 0;            // lane k < n_act: row index of open row k
        // (also with at most small64 open rows: compacting them bit by bit costs more than the
        // 32-bit passes save over the one or two passes such rows usually need)
        if (n_act > 32 || n_act <= small64) {
            bool changed = true;
            while (changed) {
                pass++;
                // per-pass copy of the group count that the compiler cannot hoist: the
                // group guards are then re-tested with scalar compares instead of being
                // kept (and spilled) as masks across the passes
                int Gp = G;
                asm volatile("" : "+s"(Gp));
                // the clause words count as rewritten every pass, so their slot addresses are
                // formed inside the pass rather than hoisted into 48 live registers
#pragma unroll
                for (int g = 0; g < kFastG; g++) asm volatile("" : "+v"(pk[g]));
                uint64_t chg = 0;       // rows in which this lane's clauses implied a literal
                // Group g's reads are issued after group g-1's implication writes (LDS
                // operations of one wavefront complete in order), so every implication is seen
                // by the groups after it in the same pass.  Reading group g+1 ahead of group g's
                // writes hides the read latency within the wavefront but misses those
                // implications until the next pass: ~8% more passes at p = 0.1, and the other
                // wavefronts of the SIMD hide the latency anyway.
                const auto fetch = [&](uint32_t (&s)[3], int g, u32x4& a, u32x4& b, u32x4& c) {
                    const uint32_t p = pk[g];
                    s[0] = slot_addr<0, 4>(dupb, p);
                    s[1] = slot_addr<1, 4>(dupb, p);
                    s[2] = slot_addr<2, 4>(dupb, p);
                    a = lds_read16(s[0]);
                    b = lds_read16(s[1]);
                    c = lds_read16(s[2]);
                };
                const auto group = [&](const uint32_t (&s)[3], const u32x4& qa, const u32x4& qb, const u32x4& qc) {
                    const uint64_t Fa = u64(qa.z, qa.w), Fb = u64(qb.z, qb.w), Fc = u64(qc.z, qc.w);
                    // per 32-bit half: T = some literal true; new conflicts and units where the
                    // row is open (neither T nor frozen), each one v_bitop3
                    const uint32_t Tl = qa.x | qb.x | qc.x, Th = qa.y | qb.y | qc.y;
                    const uint32_t fl = (uint32_t)frozen, fh = (uint32_t)(frozen >> 32);
                    const uint64_t newconf = u64(andn2(and3(qa.z, qb.z, qc.z), Tl, fl), andn2(and3(qa.w, qb.w, qc.w), Th, fh));
                    const uint64_t U = u64(andn2(two3(qa.z, qb.z, qc.z), Tl, fl), andn2(two3(qa.w, qb.w, qc.w), Th, fh));
                    if (U) {
                        // all six ORs for a lane with a unit clause (a literal that is not the
                        // unit one ORs zero): one branch instead of one per literal
                        const uint64_t Ua = U & ~Fa, Ub = U & ~Fb, Uc = U & ~Fc;
                        lds_or64(s[0], Ua);
                        lds_or64((s[0] ^ 16u) + 8u, Ua);
                        lds_or64(s[1], Ub);
                        lds_or64((s[1] ^ 16u) + 8u, Ub);
                        lds_or64(s[2], Uc);
                        lds_or64((s[2] ^ 16u) + 8u, Uc);
                    }
                    chg |= U;
                    if (ballot(newconf != 0)) {
                        frozen |= wave_or64(newconf);
                        return (frozen & amask) == amask;
                    }
                    return false;
                };
                uint32_t s3[3];
                u32x4 qa, qb, qc;
#pragma unroll
                for (int g = 0; g < kFastG; g++) {
                    if (g < Gp) {
                        fetch(s3, g, qa, qb, qc);
                        if (group(s3, qa, qb, qc)) break;
                    }
                }
                const uint64_t live = wave_or64(chg) & amask & ~frozen;   // rows still moving
                changed = live != 0ull;
                wave_sync();
                // the handoff adds latency to the wavefront (a chain of LDS round trips), so it
                // waits for the second pass and for at least 8 rows still moving: with fewer,
                // or after the first pass at low trail density, the next pass is usually the
                // last and the handoff would lengthen the launch's tail
                const int n_live = __builtin_popcountll(live);
                if (pass >= 2 && n_live >= 8 && n_live <= 32) {
                    handoff = true;
                    act = live;
                    changed = false;
                }
            }
            // complementary implications are conflicts too (nothing to look for once every
            // row is in conflict)
            if ((frozen & amask) != amask) {
                uint64_t cm = 0;
                for (int v = lane; v < 16 * W; v += kWave) {
                    const u32x4 q = dup4[2 * v];
                    cm |= u64(q.x, q.y) & u64(q.z, q.w);
                }
                frozen |= wave_or64(cm);
            }
            my_conflict = (frozen >> lane) & 1ull;
            act &= ~frozen;
            n_act = handoff ? __builtin_popcountll(act) : 0;
            if ((frozen & amask) != amask) {
                uint32_t hx[kFastW];    // lane a: word j of row a
#pragma unroll
                for (int j = 0; j < kFastW; j++) {
                    hx[j] = 0u;
                    if (j < W) {
                        hx[j] = tp(dup[4 * (32 * j + i) + half]);   // lane a: word j of row a
                        cnt += __builtin_popcount(hx[j]);
                    }
                }
                wave_sync();
                // the rows at their fixpoint are staged (the 64-bit table is dead from here on)
                const bool still_open = handoff && ((act >> lane) & 1ull);
                if (lane < A && !my_conflict && !still_open) {
#pragma unroll
                    for (int j = 0; j < kFastW; j++)
                        if (j < W) dup[lane * W + j] = hx[j];
                }
                if (n_act > 0) {
                    // lane k < n_act takes still-open row k: its index through a map in the
                    // upper half of the dead table, its words from the row's lane; then the
                    // compacted table is built here, as in the passes' own set-up below
                    uint32_t* ct = dup + 64 * W;
                    if ((act >> lane) & 1ull) ct[__builtin_popcountll(act & ((1ull << lane) - 1ull))] = (uint32_t)lane;
                    wave_sync();
                    a_k = lane < n_act ? (int)ct[lane] : 0;
                    wave_sync();
#pragma unroll
                    for (int j = 0; j < kFastW; j++) {
                        const uint32_t w = (uint32_t)__builtin_amdgcn_ds_bpermute(a_k << 2, (int)hx[j]);
                        hx[j] = lane < n_act ? w : 0u;
                    }
                    if (lane < 2) ct[64 * W + lane] = 0u;
#pragma unroll
                    for (int j = 0; j < kFastW; j++)
                        if (j < W) hx[j] = tp(hx[j]);      // lane l < 32: literal 32j+l
                    if (lane < 32) {
#pragma unroll
                        for (int j = 0; j < kFastW; j++) {
                            if (j < W) {
                                const int l = 32 * j + lane;
                                ct[2 * l] = hx[j];
                                ct[2 * (l ^ 1) + 1] = hx[j];
                            }
                        }
                    }
                }
            }
        }
        if (n_act > 0 && n_act <= 32) {
            // Few rows still open (the usual case after the evaluation already froze every
            // row with a falsified clause, or after a handoff): their bits are gathered into
            // 32-bit masks, so the passes read 8 bytes per literal and run 32-bit logic.
            // Rows are independent, so the closure of each row is unchanged.
            const bool me_act = (act >> lane) & 1ull;
            const int k_me = __builtin_popcountll(act & ((1ull << lane) - 1ull));
            uint32_t* ct = dup + 64 * W;            // slot l = {T(l), T(l ^ 1)}, then the zero slot
            const uint2* ct2 = reinterpret_cast<const uint2*>(ct);
            const uint32_t ctb = lds_addr(ct);      // 32-byte aligned: 256*W bytes past dup
            if (handoff) {
                // the compacted table was built at the handoff
            } else if (n_act <= 8) {
                // the open rows' bits in the 64-bit table are still the input's (nothing has
                // written it since it was built): compress each slot to the open rows
                int al[8];
#pragma unroll
                for (int k = 0; k < 8; k++) al[k] = 0;
                {
                    uint64_t r = act;
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        if (k < n_act) {
                            al[k] = __builtin_ctzll(r);
                            r &= r - 1ull;
                            if (lane == k) a_k = al[k];
                        }
                    }
                }
                uint32_t ctt[8], ctf[8];
#pragma unroll
                for (int it = 0; it < 8; it++) {
                    ctt[it] = ctf[it] = 0u;
                    if (it * kWave < 32 * W && lane < 32 * W - it * kWave) {
                        const u32x4 q = dup4[it * kWave + lane];
#pragma unroll
                        for (int k = 0; k < 8; k++) {
                            if (k < n_act) {
                                const int a = al[k];
                                const uint32_t tw = a < 32 ? q.x : q.y, fw = a < 32 ? q.z : q.w;
                                ctt[it] |= ((tw >> (a & 31)) & 1u) << k;
                                ctf[it] |= ((fw >> (a & 31)) & 1u) << k;
                            }
                        }
                    }
                }
                wave_sync();
#pragma unroll
                for (int it = 0; it < 8; it++) {
                    if (it * kWave < 32 * W && lane < 32 * W - it * kWave) {
                        const int v = it * kWave + lane;
                        ct[2 * v] = ctt[it];
                        ct[2 * v + 1] = ctf[it];
                    }
                }
                if (lane < 2) ct[64 * W + lane] = 0u;
            } else {
                // lane k < n_act takes open row k: its row index through an LDS map (the
                // 64-bit table is dead from here on), its words straight from the input
                wave_sync();
                if (me_act) dup[k_me] = (uint32_t)lane;
                if (lane < 2) ct[64 * W + lane] = 0u;
                wave_sync();
                a_k = lane < n_act ? (int)dup[lane] : 0;
                uint32_t x[kFastW];
#pragma unroll
                for (int j = 0; j < kFastW; j++) x[j] = (j < W && lane < n_act) ? src[a_k * W + j] : 0u;
                // all transposes first, then all table writes: the reads and writes of
                // the table do not interleave, so the transposes overlap
#pragma unroll
                for (int j = 0; j < kFastW; j++)
                    if (j < W) x[j] = tp(x[j]);      // lane l < 32: literal 32j+l
                if (lane < 32) {
#pragma unroll
                    for (int j = 0; j < kFastW; j++) {
                        if (j < W) {
                            const int l = 32 * j + lane;
                            ct[2 * l] = x[j];
                            ct[2 * (l ^ 1) + 1] = x[j];
                        }
                    }
                }
            }
            wave_sync();
            const uint32_t amask32 = n_act >= 32 ? ~0u : ((1u << n_act) - 1u);
            uint32_t fz = 0u;
            bool changed = true;
            while (changed) {
                pass++;
                // per-pass copy of the group count that the compiler cannot hoist: the
                // group guards are then re-tested with scalar compares instead of being
                // kept (and spilled) as masks across the passes
                int Gp = G;
                asm volatile("" : "+s"(Gp));
                // the clause words count as rewritten every pass, so their slot addresses are
                // formed inside the pass rather than hoisted into 48 live registers
#pragma unroll
                for (int g = 0; g < kFastG; g++) asm volatile("" : "+v"(pk[g]));
                uint32_t chg = 0;       // any implication written by this lane
                // groups in order, each read after the previous group's writes (as in the
                // 64-bit passes).  s[] = the slot addresses.
                const auto fetch = [&](uint32_t (&s)[3], int g, uint2& a, uint2& b, uint2& c) {
                    const uint32_t p = pk[g];
                    s[0] = slot_addr<0, 3>(ctb, p);
                    s[1] = slot_addr<1, 3>(ctb, p);
                    s[2] = slot_addr<2, 3>(ctb, p);
                    a = lds_read8(s[0]);
                    b = lds_read8(s[1]);
                    c = lds_read8(s[2]);
                };
                // one group: implications of its unit clauses, new conflicts; true once
                // every open row is in conflict
                const auto group = [&](const uint32_t (&s)[3], const uint2& qa, const uint2& qb, const uint2& qc) {
                    const uint32_t Fa = qa.y, Fb = qb.y, Fc = qc.y;
                    const uint32_t T = qa.x | qb.x | qc.x;
                    const uint32_t newconf = andn2(and3(Fa, Fb, Fc), T, fz);
                    const uint32_t U = andn2(two3(Fa, Fb, Fc), T, fz);
                    if (U) {
                        const uint32_t Ua = U & ~Fa, Ub = U & ~Fb, Uc = U & ~Fc;
                        lds_or32(s[0], Ua);
                        lds_or32((s[0] ^ 8u) + 4u, Ua);
                        lds_or32(s[1], Ub);
                        lds_or32((s[1] ^ 8u) + 4u, Ub);
                        lds_or32(s[2], Uc);
                        lds_or32((s[2] ^ 8u) + 4u, Uc);
                    }
                    chg |= U;
                    if (ballot(newconf != 0)) {
                        fz |= wave_or32(newconf);
                        return (fz & amask32) == amask32;
                    }
                    return false;
                };
                uint32_t s3[3];
                uint2 qa, qb, qc;
#pragma unroll
                for (int g = 0; g < kFastG; g++) {
                    if (g < Gp) {
                        fetch(s3, g, qa, qb, qc);
                        if (group(s3, qa, qb, qc)) break;
                    }
                }
                changed = ballot(chg != 0) != 0ull && (fz & amask32) != amask32;
                wave_sync();
            }
            uint32_t cm32 = fz;
            if ((fz & amask32) != amask32) {
                uint32_t c = 0;
                for (int v = lane; v < 16 * W; v += kWave) {
                    const uint2 q = ct2[2 * v];
                    c |= q.x & q.y;
                }
                cm32 |= wave_or32(c);
            }
            my_conflict = ((frozen >> lane) & 1ull) || (me_act && ((cm32 >> k_me) & 1u));
            if ((cm32 & amask32) != amask32) {
                // lane k < n_act transposes open row k back and stages it at its row's place
                uint32_t x[kFastW];
                int cnt_k = 0;
#pragma unroll
                for (int j = 0; j < kFastW; j++) {
                    x[j] = 0u;
                    if (j < W) {
                        x[j] = tp(lane < 32 ? ct[2 * (32 * j + lane)] : 0u);   // lane k: word j of open row k
                        cnt_k += __builtin_popcount(x[j]);
                    }
                }
                if (lane < n_act) {
#pragma unroll
                    for (int j = 0; j < kFastW; j++)
                        if (j < W) dup[a_k * W + j] = x[j];
                }
                const int c_open = __shfl(cnt_k, k_me);
                if (me_act) cnt = c_open;
            }
        }
        // the closure contains the input, so the implied literals are the difference of the
        // counts; the rows that gained some are staged in LDS and written again
        const bool open_row = lane < A && !my_conflict;
        const int np = open_row ? cnt - in_pop : 0;
        const uint64_t grown = ballot(np > 0);
        if (grown) {
            // the copies above were stored by other lanes: they complete before any rewrite
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            wave_sync();
            // row of word w: w * ceil(2^16 / W) >> 16 is exact for w < 1024 and W <= 16.
            // The lane index, word count and output pointer are made opaque here so that the
            // compiler does not keep the initial copy's addresses and guards alive (in
            // spilled scalar registers) through the whole kernel to reuse them.
            const uint32_t inv = (65536u + (uint32_t)W - 1u) / (uint32_t)W;
            // (the pointer goes through the asm as an integer and comes back as a global-memory
            // pointer, so the stores stay global stores rather than flat ones)
            int ln = lane, nw2 = nw;
            uint64_t dst_bits = (uint64_t)(uintptr_t)dst;
            asm volatile("" : "+v"(ln), "+s"(nw2), "+s"(dst_bits));
            __attribute__((address_space(1))) uint32_t* dst2 = (__attribute__((address_space(1))) uint32_t*)dst_bits;
#pragma unroll
            for (int u = 0; u < kFastW; u++) {
                const int w = u * kWave + ln;
                if (u * kWave < nw2 && w < nw2 && ((grown >> (((uint32_t)w * inv) >> 16)) & 1ull))
                    __builtin_nontemporal_store(dup[w], dst2 + w);
            }
        }
        if (lane < A) {
            conflict[(int64_t)inst * A + lane] = my_conflict ? 1 : 0;
            props[(int64_t)inst * A + lane] = (int32_t)np;
        }
        if (passes_out && lane == 0) passes_out[inst] = pass;
    }
}

// ---------------------------------------------------------------------------------------
// Packed general path: instances whose clauses hold at most 6 literals (any mix of lengths),
// with n <= 496 (W <= 31 words per row) and any m.  Clause c is one 8-byte word pair of the
// batch's pkk array: slot k's 10-bit literal code sits at bit 10 * (k % 3) of word k / 3.
// Slots past the clause's length name the always-false slot 32W + 1, which is false in every
// row and never implied, so a padded clause has the status and the unit literal of the
// clause itself; lanes past m name the zero slot 32W (neither true nor false), which gives
// no status and no unit.  The body runs for K = the instance's longest clause (at least 3),
// so a 3-SAT instance reads three slots per clause and a 5-SAT one five.
//
// One workgroup of NW wavefronts per instance: the waves share the instance's mask table in
// LDS (the 3-literal kernel's slot l = {T(l), T(l ^ 1)}, 16 bytes, partner slot at address
// bit 4) and take its clause groups in turn (wave v: groups v, v + NW, ...), so a long
// instance is swept NW times faster than by one wave.  Implications go straight into the
// shared table; each BCP pass ends at one barrier, where the waves' "implied something"
// rows and conflict rows are combined in LDS.  The table is built straight from the input
// rows in HBM and the closure rows go straight back.  On long instances, once at most 32 rows
// still move, the rows at their fixpoint are written out and the passes go on over a table of
// 8-byte slots {T, F} holding 32-bit masks of the moving rows, built in place of the first.
//
// LDS per workgroup (words): the table, 128W + 8; then the conflict rows (2), three
// "implied something" words of 2 (pass p uses slot p % 3: the slot the next pass ORs into
// is cleared before this pass's barrier, after every wave has read it two barriers earlier),
// and 192 count totals.
constexpr int kPkAhead = 4;

__host__ __device__ constexpr int pk_lds_words(int W) { return 128 * W + 8 + 8 + 192; }

template <int K>
struct PkGroup {
    uint2 w;            // the lane's clause word pair (slot addresses are formed again from it
                        // where an implication is written, so they are not held in registers)
    u32x4 q[K];         // slot contents {T lo, T hi, F lo, F hi}
};

// byte address of slot k of the clause word pair w in the table at tabb
template <int K>
__device__ __forceinline__ uint32_t pk_addr(const uint2& w, uint32_t tabb, int k)
{
    return tabb + 16u * __builtin_amdgcn_ubfe(k < 3 ? w.x : w.y, 10 * (k % 3), 10);
}

template <int K>
__device__ __forceinline__ void pk_fetch(uint2 w, uint32_t tabb, PkGroup<K>& gr)
{
    gr.w = w;
#pragma unroll
    for (int k = 0; k < K; k++) gr.q[k] = lds_read16(pk_addr<K>(w, tabb, k));
}

// A group's slots in the compacted table of at most 32 open rows: 8-byte slots {T, F}.
template <int K>
struct PkGroup32 {
    uint2 w;
    uint2 q[K];
};

template <int K>
__device__ __forceinline__ uint32_t pk_addr8(const uint2& w, uint32_t tabb, int k)
{
    return tabb + 8u * __builtin_amdgcn_ubfe(k < 3 ? w.x : w.y, 10 * (k % 3), 10);
}

template <int K>
__device__ __forceinline__ void pk_fetch(uint2 w, uint32_t tabb, PkGroup32<K>& gr)
{
    gr.w = w;
#pragma unroll
    for (int k = 0; k < K; k++) gr.q[k] = lds_read8(pk_addr8<K>(w, tabb, k));
}

// The same over the compacted rows: one 32-bit word each.
template <int K>
__device__ __forceinline__ void pk_status32(const PkGroup32<K>& gr, uint32_t& T, uint32_t& N0, uint32_t& N1)
{
    if (K == 3) {
        T = gr.q[0].x | gr.q[1].x | gr.q[2].x;
        N0 = and3(gr.q[0].y, gr.q[1].y, gr.q[2].y);
        N1 = two3(gr.q[0].y, gr.q[1].y, gr.q[2].y);
        return;
    }
    T = gr.q[0].x;
    N0 = gr.q[0].y;
    N1 = ~gr.q[0].y;
#pragma unroll
    for (int k = 1; k < K; k++) {
        T |= gr.q[k].x;
        N1 = bfi(gr.q[k].y, N1, N0);
        N0 &= gr.q[k].y;
    }
}

// Per 32-bit half h: T = some literal true, N0 = every literal false, N1 = exactly one
// literal not false.
template <int K>
__device__ __forceinline__ void pk_status(const PkGroup<K>& gr, uint32_t (&T)[2], uint32_t (&N0)[2],
                                          uint32_t (&N1)[2])
{
    if (K == 3) {
        T[0] = gr.q[0].x | gr.q[1].x | gr.q[2].x;
        T[1] = gr.q[0].y | gr.q[1].y | gr.q[2].y;
        N0[0] = and3(gr.q[0].z, gr.q[1].z, gr.q[2].z);
        N0[1] = and3(gr.q[0].w, gr.q[1].w, gr.q[2].w);
        N1[0] = two3(gr.q[0].z, gr.q[1].z, gr.q[2].z);
        N1[1] = two3(gr.q[0].w, gr.q[1].w, gr.q[2].w);
        return;
    }
    T[0] = gr.q[0].x;
    T[1] = gr.q[0].y;
    N0[0] = gr.q[0].z;
    N0[1] = gr.q[0].w;
    N1[0] = ~gr.q[0].z;
    N1[1] = ~gr.q[0].w;
#pragma unroll
    for (int k = 1; k < K; k++) {
        T[0] |= gr.q[k].x;
        T[1] |= gr.q[k].y;
        // exactly one non-false so far: it was before and this literal is false, or none was
        // and this one is not
        N1[0] = bfi(gr.q[k].z, N1[0], N0[0]);
        N1[1] = bfi(gr.q[k].w, N1[1], N0[1]);
        N0[0] &= gr.q[k].z;
        N0[1] &= gr.q[k].w;
    }
}

// One sweep over this wave's clause groups g0, g0 + step, ...: a group's clause words arrive
// kPkAhead of the wave's groups ahead of their use.  With PIPE (the evaluation, which writes
// nothing to the table) and 3-literal clauses, a group's slot reads are issued before the
// previous group is processed.  BCP passes read a group's slots only after the previous
// group's implication writes, so a pass sees every implication made before it in the same
// wavefront (reading ahead would miss them until the next pass: 9-10% more time at p = 0.05
// and 0.1 on mixed k-SAT).  process(group) returns true to stop the sweep.
template <int K, bool PIPE, class Grp = PkGroup<K>, class F>
__device__ __forceinline__ void pk_sweep(const uint2* __restrict__ ipk, int m, int G, int g0, int step, int lane,
                                         uint32_t tabb, uint2 zw, F&& process)
{
    const int cnt = G > g0 ? (G - g0 + step - 1) / step : 0;   // this wave's groups
    const auto grp = [&](int t) { return g0 + t * step; };
    const auto word = [&](uint2 w, int t) { return lane < m - grp(t) * kWave ? w : zw; };
    // lanes past m re-read clause m - 1 (in bounds) and are given the zero slots
    const auto load = [&](int t) -> uint2 { return ipk[grp(t) * kWave + min(lane, m - 1 - grp(t) * kWave)]; };
    uint2 pw[kPkAhead];
#pragma unroll
    for (int u = 0; u < kPkAhead; u++) pw[u] = u < cnt ? load(u) : zw;
    // 3-literal clauses in the evaluation: the next group's slot reads go out before this
    // group is processed (two register sets); longer clauses read their own group's slots (one
    // set: 70 instead of 106 VGPRs, 7 wavefronts per SIMD instead of 4)
    constexpr bool kPipe = PIPE && K <= 3;
    Grp buf[kPipe ? 2 : 1];
    if constexpr (kPipe) {
        if (cnt > 0) pk_fetch<K>(word(pw[0], 0), tabb, buf[0]);
    }
    bool stop = false;
    for (int t0 = 0; t0 < cnt && !stop; t0 += kPkAhead) {
#pragma unroll
        for (int u = 0; u < kPkAhead; u++) {
            const int t = t0 + u;
            if (t < cnt && !stop) {
                if constexpr (kPipe) {
                    if (t + 1 < cnt) pk_fetch<K>(word(pw[(u + 1) % kPkAhead], t + 1), tabb, buf[(u + 1) & 1]);
                    if (t + kPkAhead < cnt) pw[u] = load(t + kPkAhead);
                    stop = process(buf[u & 1]);
                } else {
                    pk_fetch<K>(word(pw[u], t), tabb, buf[0]);
                    if (t + kPkAhead < cnt) pw[u] = load(t + kPkAhead);
                    stop = process(buf[0]);
                }
            }
        }
    }
}

__device__ __forceinline__ void lds_or64p(uint64_t* p, uint64_t v)
{
    __hip_atomic_fetch_or((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint64_t lds_ld64(const uint64_t* p)
{
    return __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int K, bool DO_EVAL, bool DO_BCP>
__device__ __forceinline__ void eval_bcpk_body(const InstMeta& mi, int inst, int wv, int NW, int lane, uint32_t* lds,
                                               const uint2* __restrict__ pkk, const uint32_t* __restrict__ rows, int A,
                                               int32_t* __restrict__ counts, uint32_t* __restrict__ out_rows,
                                               int32_t* __restrict__ conflict, int32_t* __restrict__ props,
                                               int32_t* __restrict__ passes_out, int compact_max)
{
    const int W = __builtin_amdgcn_readfirstlane(words_per_row(mi.n));
    const int m = __builtin_amdgcn_readfirstlane(mi.m);
    const int G = (m + kWave - 1) / kWave;
    const int64_t row0 = (int64_t)A * uni64(mi.word_base);
    const uint2* ipk = pkk + uni64(mi.coff_base - inst);
    const uint32_t* src = rows + row0;
    uint32_t* tab = lds;
    const uint32_t tabb = lds_addr(tab);
    const u32x4* tab4 = reinterpret_cast<const u32x4*>(tab);
    uint64_t* sh_frozen = reinterpret_cast<uint64_t*>(lds + 128 * W + 8);
    uint64_t* sh_chg = sh_frozen + 1;                        // [3]
    uint32_t* sh_tot = lds + 128 * W + 16;                   // [64][3]
    const uint32_t zslot = 32u * (uint32_t)W;
    const uint32_t zc = zslot | (zslot << 10) | (zslot << 20);
    const uint2 zw = make_uint2(zc, zc);
    const Transposer tp(lane);
    const int half = lane >> 5, i = lane & 31;
    const int tid = wv * kWave + lane, nthreads = NW * kWave;

    // the mask table: lane a reads word j of row a, the transpose hands lane (half, i) the
    // masks of literal 32j + i over rows 32*half .. 32*half + 31 (wave v: words v, v + NW, ...)
    for (int j0 = wv; j0 < W; j0 += 8 * NW) {
        uint32_t r[8];
#pragma unroll
        for (int u = 0; u < 8; u++) r[u] = (j0 + u * NW < W && lane < A) ? src[lane * W + j0 + u * NW] : 0u;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int j = j0 + u * NW;
            if (j < W) {
                const uint32_t x = tp(r[u]);
                const int l = 32 * j + i;
                tab[4 * l + half] = x;
                tab[4 * (l ^ 1) + 2 + half] = x;
            }
        }
    }
    if (wv == 0 && lane < 8) tab[4 * zslot + lane] = lane < 6 ? 0u : ~0u;   // zero slot, then always-false slot
    for (int k = tid; k < 8 + 192; k += nthreads) lds[128 * W + 8 + k] = 0u;
    __syncthreads();

    uint64_t unsat_any = 0;
    if (DO_EVAL) {
        uint32_t acc[2][3] = {{0, 0, 0}, {0, 0, 0}};
        pk_sweep<K, true>(ipk, m, G, wv, NW, lane, tabb, zw, [&](const PkGroup<K>& gr) {
            uint32_t T[2], N0[2], N1[2];
            pk_status<K>(gr, T, N0, N1);
            // SAT|UNIT and UNSAT|UNIT per clause
            const uint64_t b0 = u64(T[0] | N1[0], T[1] | N1[1]);
            const uint64_t b1 = u64((N0[0] | N1[0]) & ~T[0], (N0[1] | N1[1]) & ~T[1]);
            count_group(b0, b1, tp, acc);
            return false;
        });
        uint32_t tot[3];
        finish_counts(acc, lane, tot);
        if (lane < A) {
#pragma unroll
            for (int s = 0; s < 3; s++) atomicAdd(&sh_tot[3 * lane + s], tot[s]);
        }
        __syncthreads();
        const bool my_unsat = lane < A && sh_tot[3 * lane + 1] != 0u;
        if (wv == 0 && lane < A) {
            int32_t* o = counts + ((int64_t)inst * A + lane) * 3;
            o[0] = (int32_t)sh_tot[3 * lane];
            o[1] = (int32_t)sh_tot[3 * lane + 1];
            o[2] = (int32_t)sh_tot[3 * lane + 2];
        }
        unsat_any = ballot(my_unsat);
    }

    if (DO_BCP) {
        const uint64_t amask = A >= 64 ? ~0ull : ((1ull << A) - 1ull);
        // contradictory inputs, and inputs that already falsify a clause, are conflicts
        {
            uint64_t cm = unsat_any;
            for (int v = tid; v < 16 * W; v += nthreads) {
                const u32x4 q = tab4[2 * v];
                cm |= u64(q.x, q.y) & u64(q.z, q.w);
            }
            cm = wave_or64(cm);
            if (lane == 0 && cm) lds_or64p(sh_frozen, cm);
        }
        __syncthreads();
        // Every wave leaves the pass loop on the same value: the pass's "moved" slot, which no
        // wave writes between the barrier that ends the pass and the next pass's barrier.  The
        // conflict rows only grow, and any snapshot of them is a valid one.
        uint64_t frozen = 0;
        int pass = 0;
        uint64_t mv = 0;            // rows the last pass moved (the same in every wave)
        for (;;) {
            pass++;
            frozen |= lds_ld64(sh_frozen);
            uint64_t chg = 0;       // rows in which this lane's clauses implied a literal
            if ((frozen & amask) != amask) pk_sweep<K, false>(ipk, m, G, wv, NW, lane, tabb, zw, [&](const PkGroup<K>& gr) {
                uint32_t T[2], N0[2], N1[2];
                pk_status<K>(gr, T, N0, N1);
                const uint32_t fl = (uint32_t)frozen, fh = (uint32_t)(frozen >> 32);
                const uint64_t newconf = u64(andn2(N0[0], T[0], fl), andn2(N0[1], T[1], fh));
                const uint64_t U = u64(andn2(N1[0], T[0], fl), andn2(N1[1], T[1], fh));
                if (U) {
                    // the unit literal is the one slot that is not false
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const uint64_t Uk = U & ~u64(gr.q[k].z, gr.q[k].w);
                        if (Uk) {
                            const uint32_t a = pk_addr<K>(gr.w, tabb, k);
                            lds_or64(a, Uk);
                            lds_or64((a ^ 16u) + 8u, Uk);
                        }
                    }
                }
                chg |= U;
                if (ballot(newconf != 0)) {
                    frozen |= wave_or64(newconf);
                    return (frozen & amask) == amask;
                }
                return false;
            });
            // rows this wave found in conflict, rows it still moved; the next pass's slot is
            // cleared before the barrier (every wave read it two barriers ago)
            const uint64_t moved = wave_or64(chg) & ~frozen;
            if (lane == 0) {
                if (frozen) lds_or64p(sh_frozen, frozen);
                if (moved) lds_or64p(&sh_chg[pass % 3], moved);
                if (wv == 0) __hip_atomic_store((unsigned long long*)&sh_chg[(pass + 1) % 3], 0ull, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            __syncthreads();
            mv = lds_ld64(&sh_chg[pass % 3]) & amask;
            if (mv == 0ull) break;
            // From the second pass on, with at most compact_max rows still moving, the passes go
            // on over a table of 32-bit masks of those rows (a handoff) when a pass costs each
            // wave at least about 4W slot reads (G K / NW >= 4 W): the handoff itself is ~2W
            // transposes on one wave, and it pays back over the next few passes only on long
            // instances (on the mixed k-SAT workload, the 5-SAT ones: -6% at p = 0.3, no
            // change at p = 0.02-0.1; handing off after the first pass, or for shorter
            // instances, cost up to 9% at p = 0.05)
            if (pass >= 2 && __builtin_popcountll(mv) <= compact_max && G * K >= 4 * W * NW) break;
        }
        // complementary implications are conflicts too
        {
            uint64_t cm = 0;
            for (int v = tid; v < 16 * W; v += nthreads) {
                const u32x4 q = tab4[2 * v];
                cm |= u64(q.x, q.y) & u64(q.z, q.w);
            }
            cm = wave_or64(cm) | frozen;
            if (lane == 0 && cm) lds_or64p(sh_frozen, cm);
        }
        __syncthreads();
        // no wave writes the conflict rows from here on: every wave reads the same ones
        frozen = lds_ld64(sh_frozen);
        const uint64_t live = mv & ~frozen;     // rows for the compacted passes
        if (live != 0ull) {
            // Handoff.  Wave 0 reads the rows back from the 64-bit table word by word (word
            // j's slots lie at table words [128j, 128j + 128)), writes the output of every row
            // at its fixpoint or in conflict, and builds the compacted table of the live rows
            // in place (word j's 8-byte slots go to [64j, 64j + 64), below every 64-bit slot
            // still to be read).  Open row k of the compacted table is row a_k.
            const int n_live = __builtin_popcountll(live);
            const uint32_t amask32 = n_live >= 32 ? ~0u : ((1u << n_live) - 1u);
            uint32_t* sh_map = sh_tot;                  // [32]: row of open row k
            uint32_t* sh_fz = sh_tot + 32;              // open rows in conflict
            uint32_t* ct = tab;
            const uint32_t ctb = tabb;
            const uint2* ct2 = reinterpret_cast<const uint2*>(ct);
            const bool me_live = (live >> lane) & 1ull;
            int a_k = 0;
            int dcnt = 0;                               // lane a: implied literals of done row a
            if (wv == 0) {
                if (me_live) sh_map[__builtin_popcountll(live & ((1ull << lane) - 1ull))] = (uint32_t)lane;
                if (lane == 0) *sh_fz = 0u;
                wave_sync();
                a_k = lane < n_live ? (int)sh_map[lane] : 0;
                const bool my_conf = (frozen >> lane) & 1ull;
                uint32_t* orow = out_rows + row0;
                for (int j = 0; j < W; j++) {
                    const uint32_t x = tp(tab[4 * (32 * j + i) + half]);   // lane a: word j of row a
                    if (lane < A && !me_live) {
                        const uint32_t in = src[lane * W + j];
                        const uint32_t o = my_conf ? in : x;
                        dcnt += __builtin_popcount(o) - __builtin_popcount(in);
                        orow[lane * W + j] = o;
                    }
                    uint32_t xk = (uint32_t)__builtin_amdgcn_ds_bpermute(a_k << 2, (int)x);
                    xk = tp(lane < n_live ? xk : 0u);                       // lane l < 32: literal 32j + l
                    if (lane < 32) {
                        const int l = 32 * j + lane;
                        ct[2 * l] = xk;
                        ct[2 * (l ^ 1) + 1] = xk;
                    }
                }
                // the zero slot, then the always-false slot
                if (lane < 4) ct[64 * W + lane] = lane < 3 ? 0u : ~0u;
            }
            __syncthreads();
            uint32_t fz = 0u;
            for (;;) {
                pass++;
                fz |= __hip_atomic_load(sh_fz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                uint32_t chg = 0;
                if ((fz & amask32) != amask32)
                    pk_sweep<K, false, PkGroup32<K>>(ipk, m, G, wv, NW, lane, ctb, zw, [&](const PkGroup32<K>& gr) {
                        uint32_t T, N0, N1;
                        pk_status32<K>(gr, T, N0, N1);
                        const uint32_t newconf = andn2(N0, T, fz);
                        const uint32_t U = andn2(N1, T, fz);
                        if (U) {
#pragma unroll
                            for (int k = 0; k < K; k++) {
                                const uint32_t Uk = U & ~gr.q[k].y;
                                if (Uk) {
                                    const uint32_t a = pk_addr8<K>(gr.w, ctb, k);
                                    lds_or32(a, Uk);
                                    lds_or32((a ^ 8u) + 4u, Uk);
                                }
                            }
                        }
                        chg |= U;
                        if (ballot(newconf != 0u)) {
                            fz |= wave_or32(newconf);
                            return (fz & amask32) == amask32;
                        }
                        return false;
                    });
                const uint32_t moved = wave_or32(chg) & ~fz;
                if (lane == 0) {
                    if (fz) lds_or32(lds_addr(sh_fz), fz);
                    if (moved) lds_or64p(&sh_chg[pass % 3], moved);
                    if (wv == 0) __hip_atomic_store((unsigned long long*)&sh_chg[(pass + 1) % 3], 0ull,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                __syncthreads();
                if ((lds_ld64(&sh_chg[pass % 3]) & amask32) == 0ull) break;
            }
            // complementary implications are conflicts too
            {
                uint32_t c = 0u;
                for (int v = tid; v < 16 * W; v += nthreads) {
                    const uint2 q = ct2[2 * v];
                    c |= q.x & q.y;
                }
                c = wave_or32(c) | fz;
                if (lane == 0 && c) lds_or32(lds_addr(sh_fz), c);
            }
            __syncthreads();
            if (wv != 0) return;
            fz = __hip_atomic_load(sh_fz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            // lane k < n_live: open row k, transposed back from the compacted table
            const bool conf_k = lane < n_live && ((fz >> lane) & 1u);
            uint32_t* orow = out_rows + row0;
            int cnt_k = 0;
            for (int j = 0; j < W; j++) {
                const uint32_t x = tp(lane < 32 ? ct[2 * (32 * j + lane)] : 0u);   // lane k: word j of open row k
                if (lane < n_live) {
                    const uint32_t in = src[a_k * W + j];
                    const uint32_t o = conf_k ? in : x;
                    cnt_k += __builtin_popcount(o) - __builtin_popcount(in);
                    orow[a_k * W + j] = o;
                }
            }
            const int k_me = __builtin_popcountll(live & ((1ull << lane) - 1ull));
            const int cnt_open = __shfl(cnt_k, k_me);
            const bool conf_open = __shfl((int)conf_k, k_me) != 0;
            if (lane < A) {
                const bool c = me_live ? conf_open : ((frozen >> lane) & 1ull) != 0ull;
                conflict[(int64_t)inst * A + lane] = c ? 1 : 0;
                props[(int64_t)inst * A + lane] = me_live ? cnt_open : dcnt;
            }
            if (passes_out && lane == 0) passes_out[inst] = pass;
            return;
        }
        if (wv != 0) return;
        const bool my_conflict = (frozen >> lane) & 1ull;
        const bool all_conflict = (frozen & amask) == amask;
        uint32_t* orow = out_rows + row0;
        int cnt = 0, in_pop = 0;
        for (int j0 = 0; j0 < W; j0 += 8) {
            uint32_t x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                x[u] = 0u;
                if (j0 + u < W && !all_conflict) x[u] = tp(tab[4 * (32 * (j0 + u) + i) + half]);   // lane a: row a
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                if (j0 + u < W && lane < A) {
                    const uint32_t in = src[lane * W + j0 + u];
                    const uint32_t o = my_conflict ? in : x[u];
                    in_pop += __builtin_popcount(in);
                    cnt += __builtin_popcount(o);
                    orow[lane * W + j0 + u] = o;
                }
            }
        }
        // the closure contains the input, so the implied literals are the difference of the counts
        if (lane < A) {
            conflict[(int64_t)inst * A + lane] = my_conflict ? 1 : 0;
            props[(int64_t)inst * A + lane] = cnt - in_pop;
        }
        if (passes_out && lane == 0) passes_out[inst] = pass;
    }
}

// Workgroup b takes instance order[b]: the batch's instances by decreasing m * K, so the
// longest sweeps start first and the short ones fill the end of the launch.
template <bool DO_EVAL, bool DO_BCP>
__global__ __launch_bounds__(256) void k_eval_bcpk(const InstMeta* __restrict__ meta, const uint2* __restrict__ pkk,
                                                    const uint8_t* __restrict__ klen, const int32_t* __restrict__ order,
                                                    const uint32_t* __restrict__ rows, int A,
                                                    int32_t* __restrict__ counts, uint32_t* __restrict__ out_rows,
                                                    int32_t* __restrict__ conflict, int32_t* __restrict__ props,
                                                    int32_t* __restrict__ passes_out, int compact_max)
{
    extern __shared__ __align__(32) uint32_t lds_all[];
    const int NW = blockDim.x / kWave;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = threadIdx.x % kWave;
    const int inst = __builtin_amdgcn_readfirstlane(order[blockIdx.x]);
    const InstMeta mi = meta[inst];
    const int K = __builtin_amdgcn_readfirstlane((int)klen[inst]);
    if (K <= 3)
        eval_bcpk_body<3, DO_EVAL, DO_BCP>(mi, inst, wv, NW, lane, lds_all, pkk, rows, A, counts, out_rows, conflict,
                                           props, passes_out, compact_max);
    else if (K == 4)
        eval_bcpk_body<4, DO_EVAL, DO_BCP>(mi, inst, wv, NW, lane, lds_all, pkk, rows, A, counts, out_rows, conflict,
                                           props, passes_out, compact_max);
    else if (K == 5)
        eval_bcpk_body<5, DO_EVAL, DO_BCP>(mi, inst, wv, NW, lane, lds_all, pkk, rows, A, counts, out_rows, conflict,
                                           props, passes_out, compact_max);
    else
        eval_bcpk_body<6, DO_EVAL, DO_BCP>(mi, inst, wv, NW, lane, lds_all, pkk, rows, A, counts, out_rows, conflict,
                                           props, passes_out, compact_max);
}

// LDS words of one wavefront's region in the general kernel (literal masks, row staging,
// conflict words; plus the clause arena when it is staged).
static int general_lds_words(int max_n) { return 128 * words_per_row(max_n) + 8; }

size_t eval_bcp_lds_bytes(int max_n) { return (size_t)general_lds_words(max_n) * 4; }

static int stage_lds_words(const DeviceBatch& b)
{
    size_t mx = 0;
    for (const InstMeta& mi : b.h_meta) {
        const size_t L = (size_t)b.h_coff[mi.coff_base + mi.m];
        mx = std::max(mx, L + (((size_t)mi.m + 4) & ~(size_t)3));
    }
    return (general_lds_words(b.max_n) + (int)mx + 3) & ~3;   // keeps every region 16-byte aligned
}

template <bool E, bool B>
static void launch_variant(const DeviceBatch& b, bool stage, int wpb, int lds_words, const uint32_t* rows, int A,
                           int32_t* counts, uint32_t* out_rows, int32_t* conflict, int32_t* props, int32_t* passes,
                           uint32_t* gtab, hipStream_t st)
{
    dim3 grid((b.n_inst + wpb - 1) / wpb), block(kWave * wpb);
    const size_t lds = (size_t)wpb * lds_words * 4;
    if (gtab)
        hipLaunchKernelGGL((k_eval_bcp<E, B, false, true>), grid, block, 0, st, b.meta, b.lits, b.coff, rows, A, counts,
                           out_rows, conflict, props, passes, b.n_inst, lds_words, gtab);
    else if (stage)
        hipLaunchKernelGGL((k_eval_bcp<E, B, true, false>), grid, block, lds, st, b.meta, b.lits, b.coff, rows, A,
                           counts, out_rows, conflict, props, passes, b.n_inst, lds_words, gtab);
    else
        hipLaunchKernelGGL((k_eval_bcp<E, B, false, false>), grid, block, lds, st, b.meta, b.lits, b.coff, rows, A,
                           counts, out_rows, conflict, props, passes, b.n_inst, lds_words, gtab);
}

// Wavefronts per workgroup (independent instances sharing a workgroup): the tuned value
// (1, 2 or 4; measured within 2% of each other at 5 wavefronts per SIMD) is clamped so the
// workgroup's LDS stays within 64 KB.
static int waves_per_block(int lds_words)
{
    int w = tuning().eval_waves_per_block;
    w = (w >= 4) ? 4 : (w >= 2 ? 2 : 1);
    while (w > 1 && (size_t)w * lds_words * 4 > 64 * 1024) w >>= 1;
    return w;
}

EvalKernel eval_kernel(const DeviceBatch& b)
{
    if (b.fast3 && tuning().eval_fast3) return EK_FAST3;
    if (b.pkk && tuning().eval_pkk) return EK_PACKED;
    const bool stage = tuning().eval_stage_arena && (size_t)stage_lds_words(b) * 4 <= 40 * 1024;
    if (!stage && (size_t)general_lds_words(b.max_n) * 4 > 64 * 1024) return EK_CSR_HBM;
    return EK_CSR;
}

size_t eval_bcp_gtab_words(const DeviceBatch& b)
{
    if (b.n_inst == 0 || eval_kernel(b) != EK_CSR_HBM) return 0;
    const int lds_words = general_lds_words(b.max_n);
    return (size_t)((lds_words + 63) & ~63) * b.n_inst;
}

hipError_t launch_eval_bcp(const DeviceBatch& b, const uint32_t* rows, int A, int32_t* counts, uint32_t* out_rows,
                           int32_t* conflict, int32_t* props, int32_t* passes, bool do_eval, bool do_bcp,
                           hipStream_t st, uint32_t* gtab_ext)
{
    if (b.n_inst == 0) return hipSuccess;
    const EvalKernel ek = eval_kernel(b);
    if (ek == EK_FAST3) {
        const int lds_words = 128 * words_per_row(b.max_n) + 8;   // 32-byte multiple
        const int wpb = waves_per_block(lds_words);
        dim3 grid((b.n_inst + wpb - 1) / wpb), block(kWave * wpb);
        const size_t lds = (size_t)wpb * lds_words * 4;
        const int small64 = tuning().eval_fast3_small64;
        if (do_eval && do_bcp)
            hipLaunchKernelGGL((k_eval_bcp3<true, true>), grid, block, lds, st, b.meta, b.pk3, rows, A, counts, out_rows,
                               conflict, props, passes, b.n_inst, lds_words, small64);
        else if (do_eval)
            hipLaunchKernelGGL((k_eval_bcp3<true, false>), grid, block, lds, st, b.meta, b.pk3, rows, A, counts,
                               out_rows, conflict, props, passes, b.n_inst, lds_words, small64);
        else
            hipLaunchKernelGGL((k_eval_bcp3<false, true>), grid, block, lds, st, b.meta, b.pk3, rows, A, counts,
                               out_rows, conflict, props, passes, b.n_inst, lds_words, small64);
        return hipGetLastError();
    }
    if (ek == EK_PACKED) {
        // one workgroup of eval_pkk_waves wavefronts per instance (the largest instance's table)
        const int nw = std::max(1, std::min(4, tuning().eval_pkk_waves));
        dim3 grid(b.n_inst), block(kWave * nw);
        const size_t lds = (size_t)pk_lds_words(words_per_row(b.max_n)) * 4;
        const uint2* pk = reinterpret_cast<const uint2*>(b.pkk);
        const int cmax = std::max(0, std::min(32, tuning().eval_pkk_compact));
        if (do_eval && do_bcp)
            hipLaunchKernelGGL((k_eval_bcpk<true, true>), grid, block, lds, st, b.meta, pk, b.klen, b.korder, rows, A,
                               counts, out_rows, conflict, props, passes, cmax);
        else if (do_eval)
            hipLaunchKernelGGL((k_eval_bcpk<true, false>), grid, block, lds, st, b.meta, pk, b.klen, b.korder, rows, A,
                               counts, out_rows, conflict, props, passes, cmax);
        else
            hipLaunchKernelGGL((k_eval_bcpk<false, true>), grid, block, lds, st, b.meta, pk, b.klen, b.korder, rows, A,
                               counts, out_rows, conflict, props, passes, cmax);
        return hipGetLastError();
    }
    // stage the arena in LDS when the largest instance keeps the wavefront under 40 KB
    const int staged = stage_lds_words(b);
    const bool stage = tuning().eval_stage_arena && (size_t)staged * 4 <= 40 * 1024;
    int lds_words = stage ? staged : general_lds_words(b.max_n);
    // mask tables beyond a workgroup's 64 KB of LDS go to an HBM slice per instance, held
    // by the batch from the first such launch on
    uint32_t* gtab = nullptr;
    if (ek == EK_CSR_HBM) {
        lds_words = (lds_words + 63) & ~63;
        const size_t need = (size_t)lds_words * b.n_inst;
        if (gtab_ext) {
            gtab = gtab_ext;
        } else if (b.gtab_words < need) {
            if (b.gtab) (void)hipFree(b.gtab);
            b.gtab = nullptr;
            b.gtab_words = 0;
            hipError_t e = hipMalloc(&b.gtab, need * 4);
            if (e != hipSuccess) return e;
            b.gtab_words = need;
        }
        if (!gtab) gtab = b.gtab;
    }
    const int wpb = gtab ? 4 : waves_per_block(lds_words);
    if (do_eval && do_bcp)
        launch_variant<true, true>(b, stage, wpb, lds_words, rows, A, counts, out_rows, conflict, props, passes, gtab, st);
    else if (do_eval)
        launch_variant<true, false>(b, stage, wpb, lds_words, rows, A, counts, out_rows, conflict, props, passes, gtab,
                                    st);
    else
        launch_variant<false, true>(b, stage, wpb, lds_words, rows, A, counts, out_rows, conflict, props, passes, gtab,
                                    st);
    return hipGetLastError();
}

}  // namespace gpsat

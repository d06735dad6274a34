// solver.hip -- CDCL on the device: one wavefront solves one job (an instance, or an
// instance plus a cube of assumed literals); a persistent grid pulls jobs from a
// device work queue until it is empty.
//
// Reference path restated in wave-parallel form (src/SATSolver/SATSolver.cu:67-218):
//   decide -> propagate -> on conflict analyse, learn, backjump -> restart.
// Strategy surfaces of the reference and where they live here:
//   BCPStrategy       (src/BCPStrategy/WatchedClausesList.cu:46-309): two watched literals,
//                     per-literal watch lists; a wave walks a list 64 entries at a time
//                     and finds units / conflicts with ballots.
//   ConflictAnalysis  (src/ConflictAnalysis/GraphAnalyzer.cu:31-138): 1-UIP over the trail
//                     (the reference's decision-scheme analysis is kept in the CPU oracle).
//   ClauseLearning    (src/ClauseLearning/LearntClausesManager.cu:16-41): learned clauses are
//                     appended to a per-job arena and watched; the DB is reduced at restarts.
//   DecisionStrategy  (src/DecisionStrategy/VSIDS.cu:77-140, src/SATSolver/DecisionMaker.cu:45-55):
//                     VSIDS with phase saving, or the reference's last-free-var/positive rule.
//   Restarts          (src/Restarts/GeometricRestartsManager.cu:3-31): geometric schedule.
//   JobsQueue         (src/SATSolver/JobsQueue.cu:10-32): atomic queue head; every cube of an
//                     instance must be refuted before the instance is UNSAT.
#include "common.hpp"
#include "gpsat_internal.hpp"
#include "solver.hpp"

namespace gpsat {

namespace {

constexpr int32_t F_DELETED = 1;
constexpr int32_t F_USED = 2;
constexpr int32_t F_LOCKED = 4;

__device__ __forceinline__ int32_t prefix_count(uint64_t m)
{
    return (int32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int32_t popc64(uint64_t m) { return __builtin_popcountll(m); }
__device__ __forceinline__ int32_t ctz64(uint64_t m) { return __builtin_ctzll(m); }

__device__ __forceinline__ int32_t wave_inclusive_scan(int32_t x)
{
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        int32_t y = __shfl_up(x, s);
        if (lane_id() >= s) x += y;
    }
    return x;
}

// Max over the wave of a 64-bit key.
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t k)
{
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        uint32_t lo = __shfl_xor((uint32_t)k, s);
        uint32_t hi = __shfl_xor((uint32_t)(k >> 32), s);
        uint64_t o = ((uint64_t)hi << 32) | lo;
        k = o > k ? o : k;
    }
    return k;
}

__device__ __forceinline__ int32_t uni(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// Orders this wavefront's memory accesses across its lanes (per-wave LDS state and, in HBM
// mode, its slice of the global workspace): waits for its outstanding memory operations,
// then keeps the compiler from moving accesses across.  The wavefronts of a workgroup are
// independent, so no workgroup barrier is involved.
__device__ __forceinline__ void wave_mem_sync()
{
    __builtin_amdgcn_s_waitcnt(0);
    wave_sync();
}

struct Ctx {
    // instance (uniform)
    int32_t n, m, inst;
    bool all3;           // original clause c is lits[3c .. 3c+3)
    const int32_t* lits;
    const int32_t* coff;
    // per-var / per-literal state (LDS)
    int8_t* lv;          // [2n] literal value: 1 true, -1 false, 0 unassigned
    int8_t* phase;       // [n]
    int32_t* seen;       // [n]
    int32_t* level;      // [n]
    int32_t* reason;     // [n]  clause id, -1 for decisions and level-0 facts
    int32_t* trail;      // [n]
    int32_t* tlim;       // [n+2] trail index where each level starts
    float* act;          // [n]
    int32_t* wstart;     // [2n+1] start of each original-clause watch list in Wo
    int32_t* wcnt_o;     // [2n]
    int32_t* wcnt_l;     // [2n]
    int32_t* lbuf;       // [n+1] clause being learned
    int32_t* lstamp;     // [n+2] level stamps for LBD
    int32_t* bumps;      // [n]
    // clause part (LDS or HBM)
    int32_t* Wo;         // original-clause watch lists (capacity = occurrences)
    int32_t* Wl;         // learned-clause watch lists, slack slots per literal
    int32_t* wpair;      // [2*m] the two watched literals of every original clause
    int32_t* wpl;        // [2*learnt_cap] the same for the learned clauses (clause m + i at 2i)
    int32_t* lmeta;      // [4*learnt_cap] start, len, flags, lbd; then [learnt_cap] remap slots
    int32_t* larena;     // learned literals
    int32_t slack, learnt_cap, larena_cap;
    // search state (uniform)
    int32_t dl, tsize, qhead;
    int32_t nlearnt, larena_used;
    float inc;
    int32_t stamp;
    int32_t base_trail;      // trail entries of the formula's unit clauses
    int32_t setup_inst;      // instance whose watch lists are built (-1: none)
    bool setup_ok;
    bool ran;            // the current job set up its instance (its per-var state is valid)
    int64_t conf_reported;  // conflicts of the current job already added to its instance pool
    uint64_t ph_t;           // telemetry: tick at which the current phase began
    int32_t ph_cur;          // telemetry: the current phase (Phase)
    unsigned long long* ph_out;  // telemetry: this job's JS_PH_* counters (null: off)
    int64_t conflicts, decisions, props, learned, restarts;
    int32_t levelmax;
};

__device__ __forceinline__ void clause_span(const Ctx& J, int32_t c, const int32_t*& cl, int32_t& len)
{
    if (c < J.m && J.all3) {
        cl = J.lits + 3 * c;
        len = 3;
    } else if (c < J.m) {
        const int32_t b = J.coff[c];
        len = J.coff[c + 1] - b;
        cl = J.lits + b;
    } else {
        const int32_t* mt = J.lmeta + 4 * (c - J.m);
        cl = J.larena + mt[0];
        len = mt[1];
    }
}

// Uniform assignment: every lane stores the same values.
__device__ __forceinline__ void assign(Ctx& J, int32_t u, int32_t r)
{
    const int32_t v = u >> 1;
    J.lv[u] = 1;
    J.lv[u ^ 1] = -1;
    J.level[v] = J.dl;
    J.reason[v] = r;
    J.trail[J.tsize] = u;
    J.tsize++;
}

// Walk the watch list of the false literal fl, 64 entries per step.  Moves watches,
// compacts the list in place, assigns units in lane order and reports the first
// conflict.  Returns the new list length.
template <bool LEARNT>
__device__ int32_t walk_list(Ctx& J, int32_t fl, int32_t* W, int32_t base, int32_t cnt, int32_t& confl)
{
    const int lane = lane_id();
    int32_t keep = 0;
    for (int32_t i0 = 0; i0 < cnt; i0 += kWave) {
        const int32_t i = i0 + lane;
        const bool active = i < cnt;
        const int32_t c = active ? W[base + i] : 0;
        bool kept = active;
        int32_t unit = -1;
        bool conf = false;
        if (active && confl < 0) {
            bool dead = false;
            if (LEARNT) dead = (J.lmeta[4 * (c - J.m) + 2] & F_DELETED) != 0;
            if (dead) {
                kept = false;
            } else {
                int32_t* wp = LEARNT ? J.wpl + 2 * (c - J.m) : J.wpair + 2 * c;
                const int32_t w0 = wp[0], w1 = wp[1];
                const int32_t other = (w0 == fl) ? w1 : w0;
                if (w0 != fl && w1 != fl) {
                    kept = false;   // stale entry: the clause moved its watch away from fl
                } else if (J.lv[other] != 1) {
                    int32_t rep = -1;
                    if (!LEARNT && J.all3) {
                        // three literals, two of them the watches: the candidate is the third
                        // (no offset reads, no search loop)
                        const int32_t* cl = J.lits + 3 * c;
                        const int32_t q = (int32_t)((uint32_t)cl[0] + (uint32_t)cl[1] + (uint32_t)cl[2] -
                                                    (uint32_t)w0 - (uint32_t)w1);
                        if (q != w0 && q != w1 && J.lv[q] != -1) rep = q;
                    } else {
                        const int32_t* cl;
                        int32_t len;
                        clause_span(J, c, cl, len);
                        for (int32_t k = 0; k < len; k++) {
                            const int32_t q = cl[k];
                            if (q != w0 && q != w1 && J.lv[q] != -1) {
                                rep = q;
                                break;
                            }
                        }
                    }
                    if (rep >= 0) {
                        kept = false;
                        if (!LEARNT) {
                            const int32_t pos = atomicAdd(&J.wcnt_o[rep], 1);
                            J.Wo[J.wstart[rep] + pos] = c;
                            wp[0] = other;
                            wp[1] = rep;
                        } else {
                            const int32_t pos = atomicAdd(&J.wcnt_l[rep], 1);
                            if (pos < J.slack) {
                                J.Wl[rep * J.slack + pos] = c;
                                wp[0] = other;
                                wp[1] = rep;
                            } else {
                                // no room: the learned clause stops propagating (it may still
                                // be a reason; its literals stay in the arena)
                                atomicSub(&J.wcnt_l[rep], 1);
                                atomicOr(&J.lmeta[4 * (c - J.m) + 2], F_DELETED);
                            }
                        }
                    } else if (J.lv[other] == -1) {
                        conf = true;
                    } else {
                        unit = other;
                    }
                }
            }
        }
        const uint64_t km = ballot(kept);
        if (kept) W[base + keep + prefix_count(km)] = c;
        keep += popc64(km);
        const uint64_t cm = ballot(conf);
        if (cm && confl < 0) confl = uni(readlane(c, ctz64(cm)));
        uint64_t um = ballot(unit >= 0);
        while (um && confl < 0) {
            const int j = ctz64(um);
            um &= um - 1;
            const int32_t u = uni(readlane(unit, j));
            const int32_t rc = uni(readlane(c, j));
            const int8_t val = J.lv[u];
            if (val == -1) confl = rc;
            else if (val == 0) assign(J, u, rc);
        }
    }
    return keep;
}

// walk_list for the original clauses of a batch whose clauses all have three literals, with
// one divergent region (the watch moves) instead of a branch per test: every lane reads its
// entry's clause, watch pair, the other watch's value and the third literal's value, and the
// outcome (keep, move, unit, conflict) is computed from them.  A clause's third literal is the
// sum of its literals minus its two watches, so no search loop either.
__device__ int32_t walk_list3(Ctx& J, int32_t fl, int32_t base, int32_t cnt, int32_t& confl)
{
    const int lane = lane_id();
    int32_t* W = J.Wo;
    int32_t keep = 0;
    for (int32_t i0 = 0; i0 < cnt; i0 += kWave) {
        const int32_t i = i0 + lane;
        const bool in_list = i < cnt;
        const bool active = in_list && confl < 0;
        // lanes past the end re-read the last entry (in bounds) and take no action
        const int32_t c = W[base + min(i, cnt - 1)];
        const int32_t w0 = J.wpair[2 * c], w1 = J.wpair[2 * c + 1];
        const int32_t* cl = J.lits + 3 * c;
        const uint32_t lsum = (uint32_t)cl[0] + (uint32_t)cl[1] + (uint32_t)cl[2];
        const int32_t other = (w0 == fl) ? w1 : w0;
        const int32_t q = (int32_t)(lsum - (uint32_t)w0 - (uint32_t)w1);
        const int8_t vo = J.lv[other];
        const int8_t vq = J.lv[q];
        const bool watched = (w0 == fl) || (w1 == fl);   // else a stale entry: dropped
        const bool open = active && watched && vo != 1;
        const bool can_move = q != w0 && q != w1 && vq != -1;
        const bool move = open && can_move;
        const bool conf = open && !can_move && vo == -1;
        const int32_t unit = (open && !can_move && vo == 0) ? other : -1;
        const bool kept = in_list && !(active && (!watched || move));
        if (move) {
            const int32_t pos = atomicAdd(&J.wcnt_o[q], 1);
            J.Wo[J.wstart[q] + pos] = c;
            J.wpair[2 * c] = other;
            J.wpair[2 * c + 1] = q;
        }
        const uint64_t km = ballot(kept);
        if (kept) W[base + keep + prefix_count(km)] = c;
        keep += popc64(km);
        const uint64_t cm = ballot(conf);
        if (cm && confl < 0) confl = uni(readlane(c, ctz64(cm)));
        uint64_t um = ballot(unit >= 0);
        while (um && confl < 0) {
            const int j = ctz64(um);
            um &= um - 1;
            const int32_t u = uni(readlane(unit, j));
            const int32_t rc = uni(readlane(c, j));
            const int8_t val = J.lv[u];
            if (val == -1) confl = rc;
            else if (val == 0) assign(J, u, rc);
        }
    }
    return keep;
}

__device__ int32_t propagate(Ctx& J)
{
    int32_t confl = -1;
    while (confl < 0 && J.qhead < J.tsize) {
        const int32_t p = uni(J.trail[J.qhead]);
        J.qhead++;
        J.props++;
        const int32_t fl = p ^ 1;
        const int32_t wb = uni(J.wstart[fl]), wc = uni(J.wcnt_o[fl]);
        int32_t k = J.all3 ? walk_list3(J, fl, wb, wc, confl) : walk_list<false>(J, fl, J.Wo, wb, wc, confl);
        J.wcnt_o[fl] = k;
        k = walk_list<true>(J, fl, J.Wl, fl * J.slack, uni(J.wcnt_l[fl]), confl);
        J.wcnt_l[fl] = k;
    }
    return confl;
}

__device__ void backtrack(Ctx& J, int32_t b, bool save_phase)
{
    if (J.dl <= b) return;
    const int32_t start = uni(J.tlim[b + 1]);
    for (int32_t t = start + lane_id(); t < J.tsize; t += kWave) {
        const int32_t u = J.trail[t];
        const int32_t v = u >> 1;
        J.lv[u] = 0;
        J.lv[u ^ 1] = 0;
        J.reason[v] = -1;
        if (save_phase) J.phase[v] = (int8_t)(u & 1);
    }
    wave_mem_sync();
    J.tsize = start;
    J.qhead = start;
    J.dl = b;
}

// 1-UIP.  Leaves the learned clause in lbuf[0..len) with the asserting literal first and
// a literal of the backjump level second.
__device__ int32_t analyze(Ctx& J, int32_t confl, int32_t& bt, int32_t& lbd)
{
    const int lane = lane_id();
    // the loop state is wave-uniform; uni() keeps it in scalar registers, so the trail scan
    // below is a scalar loop rather than an exec-masked one
    int32_t nl = 1, pathc = 0, p = -1, idx = uni(J.tsize - 1), c = confl;
    for (;;) {
        if (c >= J.m) {
            const int32_t f = J.lmeta[4 * (c - J.m) + 2];
            J.lmeta[4 * (c - J.m) + 2] = f | F_USED;
        }
        const int32_t* cl;
        int32_t len;
        clause_span(J, c, cl, len);
        for (int32_t k0 = 0; k0 < len; k0 += kWave) {
            const int32_t k = k0 + lane;
            // every lane reads (lanes past the clause re-read its last literal and take no
            // part), so the only divergent regions are the marking and its bookkeeping
            const int32_t qr = cl[min(k, len - 1)];
            const int32_t q = k < len ? qr : -1;
            const int32_t v = qr >> 1;
            const int32_t lvl = J.level[v];
            bool mark = false;
            if (q >= 0 && q != p && lvl > 0) mark = atomicOr(&J.seen[v], 1) == 0;
            const bool cur = mark && lvl == J.dl;
            if (mark) {
                J.act[v] += J.inc;
                J.bumps[v] += 1;
            }
            const uint64_t cmask = ballot(mark && cur);
            pathc += popc64(cmask);
            const uint64_t lmask = ballot(mark && !cur);
            if (mark && !cur) J.lbuf[nl + prefix_count(lmask)] = q;
            nl += popc64(lmask);
        }
        // next marked literal on the trail, walking back from idx
        bool found = false;
        while (idx >= 0) {
            const int32_t t = idx - lane;
            const int32_t tl = J.trail[max(t, 0)];
            const bool hit = t >= 0 && J.seen[tl >> 1] != 0;
            const uint64_t hm = ballot(hit);
            if (hm) {
                idx = uni(idx - ctz64(hm));
                found = true;
                break;
            }
            idx = uni(idx - kWave);
        }
        if (!found) return -1;   // no current-level literal: internal error, reported by the caller
        p = uni(J.trail[idx]);
        idx--;
        J.seen[p >> 1] = 0;
        pathc--;
        if (pathc <= 0) break;
        c = uni(J.reason[p >> 1]);
    }
    J.lbuf[0] = p ^ 1;
    wave_mem_sync();
    // Local minimization: a literal whose reason's other literals are all in the clause (or at
    // level 0) is implied by the rest and dropped.  First every literal is tested against the
    // clause as learned (drops are marked ~q, seen flags untouched), then the clause is
    // compacted and the dropped literals' seen flags cleared, so the result does not depend on
    // lane order.  Dropping several at once is sound: resolving them out in reverse trail
    // order only meets literals that stay in the clause or are resolved out earlier.
    if (nl > 1) {
        for (int32_t k = 1 + lane; k < nl; k += kWave) {
            const int32_t q = J.lbuf[k];
            const int32_t v = q >> 1;
            const int32_t r = J.reason[v];
            if (r < 0) continue;
            const int32_t* cl;
            int32_t len;
            clause_span(J, r, cl, len);
            bool drop = true;
            for (int32_t j = 0; j < len && drop; j++) {
                const int32_t vx = cl[j] >> 1;
                if (vx != v && J.seen[vx] == 0 && J.level[vx] > 0) drop = false;
            }
            if (drop) J.lbuf[k] = ~q;
        }
        wave_mem_sync();
        int32_t kept = 1;
        for (int32_t k0 = 1; k0 < nl; k0 += kWave) {
            const int32_t k = k0 + lane;
            const int32_t x = k < nl ? J.lbuf[k] : 0;
            const bool keep = k < nl && x >= 0;
            if (k < nl && x < 0) J.seen[(~x) >> 1] = 0;
            const uint64_t km = ballot(keep);
            wave_mem_sync();
            if (keep) J.lbuf[kept + prefix_count(km)] = x;
            kept += popc64(km);
        }
        nl = uni(kept);
        wave_mem_sync();
    }
    // backjump level = highest level among the other literals; LBD = distinct levels
    J.stamp++;
    uint64_t best = 0;
    int32_t distinct = 0;
    for (int32_t k = 1 + lane; k < nl; k += kWave) {
        const int32_t q = J.lbuf[k];
        const int32_t v = q >> 1;
        J.seen[v] = 0;
        const int32_t lv = J.level[v];
        const uint64_t key = ((uint64_t)(uint32_t)lv << 32) | (uint32_t)(0x7fffffff - k);
        best = key > best ? key : best;
        if (atomicExch(&J.lstamp[lv], J.stamp) != J.stamp) distinct++;
    }
    best = wave_max_u64(best);
    for (int s = 32; s > 0; s >>= 1) distinct += __shfl_xor(distinct, s);
    lbd = uni(distinct) + 1;
    if (nl > 1) {
        bt = uni((int32_t)(best >> 32));
        const int32_t bi = uni(0x7fffffff - (int32_t)(uint32_t)best);
        const int32_t t1 = J.lbuf[1], tb = J.lbuf[bi];
        wave_mem_sync();
        J.lbuf[1] = tb;
        J.lbuf[bi] = t1;
    } else {
        bt = 0;
    }
    wave_mem_sync();
    return nl;
}

// Append lbuf[0..len) as a learned clause.  Returns its clause id.
__device__ int32_t store_learned(Ctx& J, int32_t len, int32_t lbd)
{
    const int32_t id = J.nlearnt;
    const int32_t start = J.larena_used;
    for (int32_t k = lane_id(); k < len; k += kWave) J.larena[start + k] = J.lbuf[k];
    J.lmeta[4 * id + 0] = start;
    J.lmeta[4 * id + 1] = len;
    J.lmeta[4 * id + 2] = 0;
    J.lmeta[4 * id + 3] = lbd;
    J.nlearnt++;
    J.larena_used += len;
    J.learned++;
    return J.m + id;
}

// Watch a learned clause on literals a and b (uniform).
__device__ void watch_learned(Ctx& J, int32_t c, int32_t a, int32_t b)
{
    J.wpl[2 * (c - J.m)] = a;
    J.wpl[2 * (c - J.m) + 1] = b;
    const int32_t pa = uni(J.wcnt_l[a]);
    const int32_t pb = uni(J.wcnt_l[b]) + (a == b ? 1 : 0);
    if (pa < J.slack && pb < J.slack) {
        J.Wl[a * J.slack + pa] = c;
        J.wcnt_l[a] = pa + 1;
        const int32_t pb2 = uni(J.wcnt_l[b]);
        J.Wl[b * J.slack + pb2] = c;
        J.wcnt_l[b] = pb2 + 1;
    } else {
        J.lmeta[4 * (c - J.m) + 2] |= F_DELETED;
    }
}

// At level 0: attach learned clause i (an index into lmeta).  A clause satisfied at
// level 0 is dropped, a clause with one non-false literal asserts it, a falsified
// clause means the job is UNSAT (returns false).
__device__ bool attach_level0(Ctx& J, int32_t i)
{
    const int lane = lane_id();
    const int32_t st = uni(J.lmeta[4 * i]);
    const int32_t len = uni(J.lmeta[4 * i + 1]);
    int32_t nf[2] = {-1, -1};
    int32_t nnf = 0;
    bool sat = false;
    for (int32_t k0 = 0; k0 < len && nnf < 2 && !sat; k0 += kWave) {
        const int32_t k = k0 + lane;
        const int32_t q = k < len ? J.larena[st + k] : -1;
        const int8_t v = q >= 0 ? J.lv[q] : (int8_t)-1;
        if (ballot(v == 1)) sat = true;
        uint64_t um = ballot(v == 0);
        while (um && nnf < 2) {
            const int j = ctz64(um);
            um &= um - 1;
            nf[nnf++] = uni(readlane(q, j));
        }
    }
    bool ok = true;
    if (sat) {
        J.lmeta[4 * i + 2] = F_DELETED;
    } else if (nnf == 0) {
        ok = false;
    } else if (nnf == 1) {
        J.lmeta[4 * i + 2] = F_DELETED;
        if (J.lv[nf[0]] == 0) assign(J, nf[0], -1);
    } else {
        watch_learned(J, J.m + i, nf[0], nf[1]);
    }
    wave_mem_sync();
    return ok;
}

// Reduce the learned DB at the current level.  Survivors: reasons of assigned literals
// (always), glue clauses (LBD <= 3) and, in mode 0, clauses used since the last
// reduction; mode 2 keeps the reasons only.  Clause ids are renumbered, reasons are
// remapped and the learned watch lists are rebuilt from the watch pairs, which keeps
// the two-watched-literal invariant at any level.
__device__ void reduce_db(Ctx& J, int mode)
{
    const int lane = lane_id();
    const int32_t m = J.m;
    int32_t* remap = J.lmeta + 4 * J.learnt_cap;
    for (int32_t t = lane; t < J.tsize; t += kWave) {
        const int32_t v = J.trail[t] >> 1;
        const int32_t r = J.reason[v];
        if (r >= m && J.level[v] > 0) atomicOr(&J.lmeta[4 * (r - m) + 2], F_LOCKED);   // level-0 reasons are never read
    }
    wave_mem_sync();
    int32_t ns = 0;
    for (int32_t i0 = 0; i0 < J.nlearnt; i0 += kWave) {
        const int32_t i = i0 + lane;
        bool keep = false, del = false;
        int32_t st = 0, len = 0, lbd = 0, w0 = 0, w1 = 0;
        if (i < J.nlearnt) {
            st = J.lmeta[4 * i];
            len = J.lmeta[4 * i + 1];
            const int32_t fl = J.lmeta[4 * i + 2];
            lbd = J.lmeta[4 * i + 3];
            w0 = J.wpl[2 * i];
            w1 = J.wpl[2 * i + 1];
            del = (fl & F_DELETED) != 0;
            const bool wanted = mode == 0 ? (lbd <= 3 || (fl & F_USED)) : mode == 1 ? lbd <= 3 : false;
            keep = (fl & F_LOCKED) || (!del && wanted);
        }
        const uint64_t km = ballot(keep);
        wave_mem_sync();
        const int32_t ni = ns + prefix_count(km);
        if (i < J.nlearnt) remap[i] = keep ? ni : -1;
        if (keep) {
            J.lmeta[4 * ni] = st;
            J.lmeta[4 * ni + 1] = len;
            J.lmeta[4 * ni + 2] = del ? F_DELETED : 0;
            J.lmeta[4 * ni + 3] = lbd;
            J.wpl[2 * ni] = w0;
            J.wpl[2 * ni + 1] = w1;
        }
        ns += popc64(km);
        wave_mem_sync();
    }
    // compact literals (forward copy; a destination never lies above its source)
    int32_t pos = 0;
    for (int32_t i = 0; i < ns; i++) {
        const int32_t st = uni(J.lmeta[4 * i]);
        const int32_t len = uni(J.lmeta[4 * i + 1]);
        for (int32_t k0 = 0; k0 < len; k0 += kWave) {
            const int32_t k = k0 + lane;
            const int32_t x = k < len ? J.larena[st + k] : 0;
            wave_mem_sync();
            if (k < len) J.larena[pos + k] = x;
            wave_mem_sync();
        }
        J.lmeta[4 * i] = pos;
        pos += len;
    }
    for (int32_t t = lane; t < J.tsize; t += kWave) {
        const int32_t v = J.trail[t] >> 1;
        const int32_t r = J.reason[v];
        if (r >= m) J.reason[v] = remap[r - m] >= 0 ? m + remap[r - m] : -1;
    }
    for (int32_t l = lane; l < 2 * J.n; l += kWave) J.wcnt_l[l] = 0;
    wave_mem_sync();
    for (int32_t i = lane; i < ns; i += kWave) {
        if (J.lmeta[4 * i + 2] & F_DELETED) continue;
        const int32_t c = m + i;
        const int32_t wl[2] = {J.wpl[2 * i], J.wpl[2 * i + 1]};
        for (int h = 0; h < 2; h++) {
            const int32_t p = atomicAdd(&J.wcnt_l[wl[h]], 1);
            if (p < J.slack) {
                J.Wl[wl[h] * J.slack + p] = c;
            } else {
                atomicSub(&J.wcnt_l[wl[h]], 1);
                J.lmeta[4 * i + 2] = F_DELETED;
            }
        }
    }
    J.nlearnt = ns;
    J.larena_used = pos;
    wave_mem_sync();
}

__device__ int32_t pick_branch(Ctx& J, int mode)
{
    uint64_t best = 0;
    for (int32_t v = lane_id(); v < J.n; v += kWave) {
        if (J.lv[2 * v] == 0) {
            uint32_t hi = mode == 1 ? 1u : (__float_as_uint(J.act[v]) | 0x80000000u) ;
            // larger key wins: activity first (non-negative floats order as integers),
            // then the lower var (VSIDS) or the higher var (reference: last free var)
            uint32_t lo = mode == 1 ? (uint32_t)v + 1u : 0xffffffffu - (uint32_t)v;
            uint64_t key = ((uint64_t)hi << 32) | lo;
            best = key > best ? key : best;
        }
    }
    best = wave_max_u64(best);
    if (best == 0) return -1;
    const uint32_t lo = (uint32_t)best;
    return uni(mode == 1 ? (int32_t)(lo - 1u) : (int32_t)(0xffffffffu - lo));
}

__device__ void rescale_activity(Ctx& J)
{
    for (int32_t v = lane_id(); v < J.n; v += kWave) J.act[v] *= 1e-30f;
    J.inc *= 1e-30f;
    wave_mem_sync();
}

__device__ __forceinline__ int32_t load_status(int32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lay out the per-wave arrays.  The state part always sits in LDS; the clause part
// sits in LDS (kLds) or in this wave's slice of the HBM workspace.
struct Layout {
    static __host__ __device__ size_t state_bytes(int32_t n)
    {
        size_t b = 0;
        b += ((2 * (size_t)n + 3) & ~(size_t)3);        // lv
        b += (((size_t)n + 3) & ~(size_t)3);            // phase
        b += 4 * (size_t)n * 5;                         // seen level reason trail act
        b += 4 * ((size_t)n + 2);                       // tlim
        b += 4 * (2 * (size_t)n + 1);                   // wstart
        b += 4 * (2 * (size_t)n) * 2;                   // wcnt_o wcnt_l
        b += 4 * ((size_t)n + 1);                       // lbuf
        b += 4 * ((size_t)n + 2);                       // lstamp
        b += 4 * (size_t)n;                             // bumps
        return (b + 15) & ~(size_t)15;
    }
};

__device__ void carve_state(Ctx& J, char* s, int32_t n)
{
    J.lv = (int8_t*)s;                 s += ((2 * n + 3) & ~3);
    J.phase = (int8_t*)s;              s += ((n + 3) & ~3);
    J.seen = (int32_t*)s;              s += 4 * n;
    J.level = (int32_t*)s;             s += 4 * n;
    J.reason = (int32_t*)s;            s += 4 * n;
    J.trail = (int32_t*)s;             s += 4 * n;
    J.act = (float*)s;                 s += 4 * n;
    J.tlim = (int32_t*)s;              s += 4 * (n + 2);
    J.wstart = (int32_t*)s;            s += 4 * (2 * n + 1);
    J.wcnt_o = (int32_t*)s;            s += 4 * (2 * n);
    J.wcnt_l = (int32_t*)s;            s += 4 * (2 * n);
    J.lbuf = (int32_t*)s;              s += 4 * (n + 1);
    J.lstamp = (int32_t*)s;            s += 4 * (n + 2);
    J.bumps = (int32_t*)s;
}

__device__ void carve_clauses(Ctx& J, int32_t* c, const SolveArgs& a)
{
    J.Wo = c;      c += a.wo_cap;
    J.Wl = c;      c += a.wl_cap;
    J.wpair = c;
    J.wpl = c + (a.wpair_cap - 2 * a.o.learnt_cap);
    c += a.wpair_cap;
    J.lmeta = c;   c += a.lmeta_cap;
    J.larena = c;
}

// Build the per-instance part of a job: watch lists of the original clauses and the
// formula's unit clauses at level 0.  Returns false if the formula is falsified there.
__device__ bool setup_instance(Ctx& J, const SolveArgs& a, int32_t inst)
{
    const int lane = lane_id();
    const SolveOptsDev& o = a.o;
    const InstMeta mi = a.meta[inst];
    J.inst = inst;
    J.n = uni(mi.n);
    J.m = uni(mi.m);
    J.lits = a.lits + mi.lit_base;
    J.coff = a.coff + mi.coff_base;
    J.all3 = a.all3 != 0;
    const int32_t n = J.n, m = J.m, n2 = 2 * n;
    const int64_t vb = a.var_base[inst];
    J.slack = o.watch_slack;
    J.learnt_cap = o.learnt_cap;
    J.larena_cap = o.larena_cap;
    J.dl = 0;
    J.tsize = 0;
    J.qhead = 0;
    J.nlearnt = 0;
    J.larena_used = 0;
    J.inc = 1.0f;
    J.stamp = 0;
    for (int32_t v = lane; v < n; v += kWave) {
        J.seen[v] = 0;
        J.level[v] = 0;
        J.reason[v] = -1;
        J.act[v] = a.act_init ? a.act_init[vb + v] : 0.0f;
        J.phase[v] = a.phase_init ? a.phase_init[vb + v] : (int8_t)o.phase_init;
        J.bumps[v] = 0;
    }
    for (int32_t v = lane; v < n + 2; v += kWave) J.lstamp[v] = 0;
    for (int32_t l = lane; l < n2; l += kWave) {
        J.lv[l] = 0;
        J.wcnt_o[l] = 0;
        J.wcnt_l[l] = 0;
    }
    wave_mem_sync();
    // occurrence counts give the capacity of every original-clause watch list
    const int32_t L = uni(J.coff[m]);
    for (int32_t k = lane; k < L; k += kWave) atomicAdd(&J.wcnt_o[J.lits[k]], 1);
    wave_mem_sync();
    int32_t carry = 0;
    for (int32_t l0 = 0; l0 < n2; l0 += kWave) {
        const int32_t l = l0 + lane;
        const int32_t x = l < n2 ? J.wcnt_o[l] : 0;
        const int32_t inc = wave_inclusive_scan(x);
        if (l < n2) {
            J.wstart[l] = carry + inc - x;
            J.wcnt_o[l] = 0;
        }
        carry += uni(__shfl(inc, 63));
    }
    if (lane == 0) J.wstart[n2] = carry;
    wave_mem_sync();
    // watch the first two literals of every clause; units and empty clauses at level 0
    bool unsat = false;
    for (int32_t c0 = 0; c0 < m && !unsat; c0 += kWave) {
        const int32_t c = c0 + lane;
        int32_t len = -1, l0 = -1;
        if (c < m) {
            const int32_t b = J.coff[c];
            len = J.coff[c + 1] - b;
            if (len >= 1) {
                l0 = J.lits[b];
                // the two watches are distinct literals; a clause repeating one literal is a unit
                int32_t l1 = -1;
                for (int32_t k = 1; k < len; k++) {
                    const int32_t q = J.lits[b + k];
                    if (q != l0) {
                        l1 = q;
                        break;
                    }
                }
                if (l1 >= 0) {
                    len = 2;
                    J.wpair[2 * c] = l0;
                    J.wpair[2 * c + 1] = l1;
                    J.Wo[J.wstart[l0] + atomicAdd(&J.wcnt_o[l0], 1)] = c;
                    J.Wo[J.wstart[l1] + atomicAdd(&J.wcnt_o[l1], 1)] = c;
                } else {
                    len = 1;
                }
            }
        }
        if (ballot(len == 0)) unsat = true;
        uint64_t um = ballot(len == 1);
        while (um && !unsat) {
            const int j = ctz64(um);
            um &= um - 1;
            const int32_t u = uni(readlane(l0, j));
            const int8_t val = J.lv[u];
            if (val == -1) unsat = true;
            else if (val == 0) assign(J, u, -1);
        }
    }
    J.base_trail = J.tsize;
    wave_mem_sync();
    return !unsat;
}

// Next cube of the instance already set up on this wavefront: undo everything above the
// formula's unit clauses and forget the learned clauses (they may rest on the previous
// cube).  Every remaining watch pair is repaired by propagating the units again from the
// start of the trail.  Activities and saved phases carry over.
__device__ void reset_instance(Ctx& J)
{
    const int lane = lane_id();
    for (int32_t t = J.base_trail + lane; t < J.tsize; t += kWave) {
        const int32_t u = J.trail[t];
        J.lv[u] = 0;
        J.lv[u ^ 1] = 0;
        J.reason[u >> 1] = -1;
    }
    for (int32_t l = lane; l < 2 * J.n; l += kWave) J.wcnt_l[l] = 0;
    for (int32_t v = lane; v < J.n; v += kWave) J.bumps[v] = 0;
    wave_mem_sync();
    J.tsize = J.base_trail;
    J.qhead = 0;
    J.dl = 0;
    J.nlearnt = 0;
    J.larena_used = 0;
}

// Telemetry: charge the time since the last switch to the current phase and enter
// `next`.  The counters are the job's JS_PH_* fields, added to without waiting.
__device__ __forceinline__ void phase(Ctx& J, int32_t next)
{
    if (!J.ph_out) return;
    const uint64_t now = wall_clock64();
    if (lane_id() == 0) atomicAdd(&J.ph_out[J.ph_cur], (unsigned long long)(now - J.ph_t));
    J.ph_t = now;
    J.ph_cur = next;
}

// Adds the conflicts of this job not yet reported to its instance's pool; returns the
// pool's new total (wave-uniform).
__device__ int64_t report_conflicts(Ctx& J, const SolveArgs& a, int32_t inst)
{
    unsigned long long tot = 0;
    const unsigned long long delta = (unsigned long long)(J.conflicts - J.conf_reported);
    if (lane_id() == 0) tot = atomicAdd(&a.inst_conf[inst], delta) + delta;
    J.conf_reported = J.conflicts;
    return (int64_t)__shfl(tot, 0);
}

// Testing aid: append the job's held learned clauses and its level-0 literals (as unit
// clauses) to the dump as one block [inst, slots, then per clause its literals and -1].  The
// block's space is reserved only if all of it fits, so the dump stays parseable.
__device__ void dump_learned(Ctx& J, const SolveArgs& a, int32_t inst)
{
    const int lane = lane_id();
    int32_t s = 0;
    for (int32_t i = lane; i < J.nlearnt; i += kWave) s += J.lmeta[4 * i + 1] + 1;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    const int32_t n0 = J.dl > 0 ? uni(J.tlim[1]) : J.tsize;   // level-0 part of the trail
    const int64_t slots = (int64_t)uni(s) + 2 * (int64_t)n0;
    const unsigned long long total = (unsigned long long)(slots + 2);
    long long base = -1;
    if (lane == 0) {
        atomicAdd(a.learned_need, total);
        unsigned long long old = __hip_atomic_load(a.learned_pos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (old + total <= (unsigned long long)a.learned_cap) {
            const unsigned long long prev = atomicCAS(a.learned_pos, old, old + total);
            if (prev == old) {
                base = (long long)old;
                break;
            }
            old = prev;
        }
    }
    base = __shfl(base, 0);
    if (base < 0) return;
    int32_t* out = a.learned_out + base;
    if (lane == 0) {
        out[0] = inst;
        out[1] = (int32_t)slots;
    }
    int64_t p = 2;
    for (int32_t i = 0; i < J.nlearnt; i++) {
        const int32_t st = uni(J.lmeta[4 * i]), len = uni(J.lmeta[4 * i + 1]);
        for (int32_t k = lane; k < len; k += kWave) out[p + k] = J.larena[st + k];
        if (lane == 0) out[p + len] = -1;
        p += len + 1;
    }
    for (int32_t t = lane; t < n0; t += kWave) {
        out[p + 2 * t] = J.trail[t];
        out[p + 2 * t + 1] = -1;
    }
}

// Returns the job verdict: 0 SAT, 1 UNSAT, 2 UNDEF (budget), 3 skipped (instance decided).
// deadline: wall-clock tick after which the job stops as UNDEF (0: none).
__device__ int32_t run_job(Ctx& J, const SolveArgs& a, const JobDesc& jd, int64_t max_conflicts, uint64_t deadline)
{
    const SolveOptsDev& o = a.o;
    J.conflicts = J.decisions = J.props = J.learned = J.restarts = 0;
    J.levelmax = 0;
    J.ran = false;
    J.conf_reported = 0;
    if (load_status(&a.inst_status[jd.inst]) != 2) return 3;
    if (o.inst_conflicts > 0 && (int64_t)__atomic_load_n(&a.inst_conf[jd.inst], __ATOMIC_RELAXED) >= o.inst_conflicts)
        return 2;
    const uint64_t t_start = wall_clock64();
    J.ph_t = t_start;
    J.ph_cur = PH_SETUP;
    if (J.setup_inst == jd.inst) {
        reset_instance(J);
    } else {
        J.setup_ok = setup_instance(J, a, jd.inst);
        J.setup_inst = jd.inst;
    }
    J.ran = true;
    if (!J.setup_ok) return 1;
    const int32_t m = J.m;
    bool unsat = false;
    // the cube's literals are facts of this job
    for (int32_t k = 0; k < jd.cube_len && !unsat; k++) {
        const int32_t u = uni(a.cube_lits[jd.cube_off + k]);
        const int8_t val = J.lv[u];
        if (val == -1) unsat = true;
        else if (val == 0) assign(J, u, -1);
    }
    wave_mem_sync();
    if (unsat) return 1;
    phase(J, PH_BCP);
    if (propagate(J) >= 0) return 1;

    int32_t restart_thr = o.restart_first, restart_cnt = 0;
    const float decay_inv = 1.0f / o.var_decay;
    for (;;) {
        phase(J, PH_BCP);
        const int32_t confl = propagate(J);
        if (confl >= 0) {
            phase(J, PH_ANALYZE);
            J.conflicts++;
            restart_cnt++;
            if (J.dl == 0) return 1;
            int32_t bt, lbd;
            const int32_t nl = analyze(J, confl, bt, lbd);
            if (nl < 0) return 4;
            phase(J, PH_BACKTRACK);
            if (nl == 1) {
                backtrack(J, 0, o.decision_mode == 0);
                assign(J, uni(J.lbuf[0]), -1);
            } else {
                backtrack(J, bt, o.decision_mode == 0);
                phase(J, PH_ANALYZE);
                if (!(J.nlearnt < J.learnt_cap && J.larena_used + nl <= J.larena_cap)) {
                    reduce_db(J, 0);
                    if (!(J.nlearnt < J.learnt_cap && J.larena_used + nl <= J.larena_cap)) reduce_db(J, 2);
                }
                if (J.nlearnt < J.learnt_cap && J.larena_used + nl <= J.larena_cap) {
                    const int32_t cid = store_learned(J, nl, lbd);
                    wave_mem_sync();
                    watch_learned(J, cid, uni(J.lbuf[0]), uni(J.lbuf[1]));
                    assign(J, uni(J.lbuf[0]), cid);
                } else {
                    // every held clause is a reason: restart, then add the clause at level 0
                    backtrack(J, 0, o.decision_mode == 0);
                    J.restarts++;
                    restart_cnt = 0;
                    reduce_db(J, 2);
                    const int32_t cid = store_learned(J, nl, lbd);
                    wave_mem_sync();
                    if (!attach_level0(J, cid - m)) return 1;
                }
            }
            wave_mem_sync();
            J.inc *= decay_inv;
            if (J.inc > 1e30f) rescale_activity(J);
            if (max_conflicts > 0 && J.conflicts >= max_conflicts) return 2;
            if ((J.conflicts & 63) == 0) {
                const uint64_t now = wall_clock64();
                if (o.max_ticks > 0 && (int64_t)(now - t_start) > o.max_ticks) return 2;
                if (deadline && now > deadline) return 2;
                if (o.inst_conflicts > 0 && report_conflicts(J, a, jd.inst) >= o.inst_conflicts) return 2;
            }
        } else {
            if (restart_cnt >= restart_thr) {
                phase(J, PH_BACKTRACK);
                restart_thr = (int32_t)((float)restart_thr * o.restart_factor);
                restart_cnt = 0;
                backtrack(J, 0, o.decision_mode == 0);
                J.restarts++;
                if (load_status(&a.inst_status[jd.inst]) != 2) return 3;
                if (J.nlearnt * 2 >= J.learnt_cap || J.larena_used * 2 >= J.larena_cap) reduce_db(J, 0);
                continue;
            }
            phase(J, PH_DECIDE);
            const int32_t v = pick_branch(J, o.decision_mode);
            if (v < 0) return 0;
            J.decisions++;
            if (o.max_decisions > 0 && J.decisions > o.max_decisions) return 2;
            if ((J.decisions & 1023) == 0) {
                if (load_status(&a.inst_status[jd.inst]) != 2) return 3;
                const uint64_t now = wall_clock64();
                if (o.max_ticks > 0 && (int64_t)(now - t_start) > o.max_ticks) return 2;
                if (deadline && now > deadline) return 2;
            }
            J.dl++;
            if (J.dl > J.levelmax) J.levelmax = J.dl;
            J.tlim[J.dl] = J.tsize;
            const int32_t ph = o.decision_mode == 1 ? 1 : (int32_t)J.phase[v];
            assign(J, 2 * v + ph, -1);
        }
    }
}

}  // namespace

// kMode: where a job's state sits -- 1 all in LDS, 2 the per-var state in LDS and the
// clause part in this wavefront's slice of the HBM workspace, 3 the per-var state and the
// watch lists and watch pairs of the original clauses in LDS (read on every propagation
// step), the learned clauses' lists, pairs, metadata and literals in HBM, 4 everything in
// the HBM slice (formulas whose per-var state does not fit a workgroup's LDS).
template <int kMode>
__global__ __launch_bounds__(kSolveWavesPerBlock * kWave) __attribute__((amdgpu_waves_per_eu(kSolveWavesPerEu, kSolveWavesPerEu)))
void k_solve(SolveArgs a)
{
    extern __shared__ __align__(16) char smem_all[];
    const int lane = lane_id();
    // the wavefronts of a workgroup are independent (separate jobs, private LDS regions,
    // wave-level ordering only)
    const int wslot = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int64_t gwave = (int64_t)blockIdx.x * (blockDim.x / kWave) + wslot;
    char* smem = smem_all + (size_t)wslot * a.lds_per_wave;
    Ctx J;
    int32_t* cpart;
    if (kMode == 4) {
        char* slice = (char*)(a.ws + gwave * a.ws_ints_per_wave);
        carve_state(J, slice, a.max_n);
        cpart = (int32_t*)(slice + Layout::state_bytes(a.max_n));
    } else {
        carve_state(J, smem, a.max_n);
        cpart = kMode == 1 ? (int32_t*)(smem + Layout::state_bytes(a.max_n)) : a.ws + gwave * a.ws_ints_per_wave;
    }
    carve_clauses(J, cpart, a);
    if (kMode == 3) {
        int32_t* l = (int32_t*)(smem + Layout::state_bytes(a.max_n));
        J.Wo = l;
        J.wpair = l + a.wo_cap;
    }
    J.setup_inst = -1;
    J.setup_ok = false;
    J.n = 0;
    int32_t n_jobs = a.n_jobs;
    const int32_t* order = nullptr;
    int64_t budget = a.o.max_conflicts;
    if (a.epoch) {
        const int32_t e = uni(*a.epoch);
        n_jobs = uni(a.n_active[e & 1]);
        order = a.active + (int64_t)(e & 1) * a.n_jobs;
        budget = a.epoch_budget[e];
    }
    // the whole-solve deadline counts from the first wavefront's start (kept across epochs)
    uint64_t deadline = 0;
    if (a.o.deadline_ticks > 0) {
        unsigned long long t0 = 0;
        if (lane == 0) {
            const unsigned long long now = wall_clock64();
            const unsigned long long old = atomicCAS(a.t_start, 0ull, now);
            t0 = old ? old : now;
        }
        t0 = __shfl(t0, 0);
        deadline = (uint64_t)t0 + (uint64_t)a.o.deadline_ticks;
    }
    int32_t j = 0, j_end = 0;
    for (;;) {
        if (j >= j_end) {
            // grab the next chunk of consecutive jobs (cubes of one instance are adjacent)
            int32_t k = 0;
            if (lane == 0) k = atomicAdd(a.queue_head, 1);
            k = uni(__shfl(k, 0));
            j = k * a.chunk;
            if (j >= n_jobs) break;
            j_end = min(j + a.chunk, n_jobs);
        }
        const JobDesc jd = a.jobs[order ? order[j] : j];
        j++;
        const uint64_t t_job = wall_clock64();
        // past the deadline a job is not started: UNDEF
        J.ran = false;
        J.conflicts = J.decisions = J.props = J.learned = J.restarts = 0;
        J.levelmax = 0;
        J.ph_out = (a.o.telemetry && a.job_stats)
                       ? (unsigned long long*)(a.job_stats + (int64_t)jd.order * JS_COUNT + JS_PH_SETUP) : nullptr;
        J.ph_t = t_job;
        J.ph_cur = PH_SETUP;
        // in epoch mode the epoch's budget replaces the per-job one, bounded by what is left
        // of opts.max_conflicts after the job's earlier epochs (a spent job is UNDEF)
        int64_t job_budget = budget;
        if (a.epoch && a.o.max_conflicts > 0 && a.job_stats) {
            const int64_t left = a.o.max_conflicts - a.job_stats[(int64_t)jd.order * JS_COUNT + JS_CONFLICTS];
            job_budget = min(budget, left);
        }
        const bool spent = a.epoch && a.o.max_conflicts > 0 && a.job_stats && job_budget <= 0;
        const int32_t st = ((deadline && t_job > deadline) || spent) ? 2 : run_job(J, a, jd, job_budget, deadline);
        phase(J, PH_SETUP);   // charge the last phase
        const uint64_t job_ticks = wall_clock64() - t_job;
        wave_mem_sync();
        const int32_t inst = jd.inst;
        if (a.o.inst_conflicts > 0 && J.ran) (void)report_conflicts(J, a, inst);
        if (a.learned_out && J.ran && J.setup_ok) dump_learned(J, a, inst);
        if (st == 0) {
            int32_t won = 0;
            if (lane == 0) won = atomicCAS(&a.inst_status[inst], 2, 0) == 2;
            won = uni(__shfl(won, 0));
            if (won) {
                const InstMeta mi = a.meta[inst];
                const int32_t W = words_per_row(mi.n);
                for (int32_t w = lane; w < W; w += kWave) {
                    uint32_t x = 0;
                    for (int b = 0; b < 32; b++) {
                        const int32_t l = 32 * w + b;
                        if (l < 2 * mi.n && J.lv[l] == 1) x |= 1u << b;
                    }
                    a.model_rows[mi.word_base + w] = x;
                }
            }
        } else if (st == 1) {
            if (lane == 0) {
                const int32_t r = atomicAdd(&a.inst_refuted[inst], 1) + 1;
                if (r == a.inst_ncubes[inst]) atomicCAS(&a.inst_status[inst], 2, 1);
            }
        } else if (st == 2) {
            if (lane == 0) atomicAdd(&a.inst_undef[inst], 1);
            if (a.act_store && J.ran) {
                // a job stopped at an epoch boundary restarts next epoch from these
                const int64_t vb = a.var_base[inst];
                for (int32_t v = lane; v < J.n; v += kWave) {
                    a.act_store[vb + v] = J.act[v];
                    a.phase_store[vb + v] = J.phase[v];
                }
            }
        }
        // per-var state belongs to this job's instance only if the job set it up (a job
        // skipped because its instance was decided, or because the deadline passed, did not)
        if (a.var_bumps && J.ran) {
            const int64_t vb = a.var_base[inst];
            for (int32_t v = lane; v < J.n; v += kWave)
                if (J.bumps[v]) atomicAdd(&a.var_bumps[vb + v], J.bumps[v]);
        }
        if (a.job_stats && lane == 0) {
            int64_t* s = a.job_stats + (int64_t)jd.order * JS_COUNT;
            // counters add up over the epochs a job runs in
            s[JS_STATUS] = st;
            s[JS_CONFLICTS] += J.conflicts;
            s[JS_DECISIONS] += J.decisions;
            s[JS_PROPS] += J.props;
            s[JS_LEARNED] += J.learned;
            s[JS_RESTARTS] += J.restarts;
            s[JS_LEVELMAX] = max(s[JS_LEVELMAX], (int64_t)J.levelmax);
            s[JS_TICKS] += (int64_t)job_ticks;
        }
        wave_mem_sync();
    }
}

// Restart epochs.  begin: empty the queue and the next job list.  end: every job of
// this epoch stopped by the budget whose instance is still open goes to the next list;
// then the epoch counter advances and the number of open jobs is published.
__global__ void k_epoch_begin(int32_t* queue_head, const int32_t* epoch, int32_t* n_active)
{
    if (threadIdx.x == 0) {
        *queue_head = 0;
        n_active[(*epoch + 1) & 1] = 0;
    }
}

__global__ void k_epoch_end(SolveArgs a, int32_t* active, int32_t* n_active)
{
    const int32_t e = *a.epoch;
    const int32_t n = n_active[e & 1];
    const int32_t* cur = active + (int64_t)(e & 1) * a.n_jobs;
    int32_t* next = active + (int64_t)((e + 1) & 1) * a.n_jobs;
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int32_t job = cur[k];
        const JobDesc jd = a.jobs[job];
        const int64_t* js = a.job_stats + (int64_t)jd.order * JS_COUNT;
        const bool spent = a.o.max_conflicts > 0 && js[JS_CONFLICTS] >= a.o.max_conflicts;
        if (js[JS_STATUS] == 2 && !spent && a.inst_status[jd.inst] == 2) next[atomicAdd(&n_active[(e + 1) & 1], 1)] = job;
    }
}

__global__ void k_epoch_advance(int32_t* epoch, const int32_t* n_active, int32_t* remaining)
{
    if (threadIdx.x == 0) {
        const int32_t e = *epoch;
        *remaining = n_active[(e + 1) & 1];
        *epoch = e + 1;
    }
}

hipError_t launch_epoch_begin(const SolveArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL(k_epoch_begin, dim3(1), dim3(64), 0, st, a.queue_head, a.epoch, (int32_t*)a.n_active);
    return hipGetLastError();
}

hipError_t launch_epoch_end(const SolveArgs& a, int32_t* remaining, hipStream_t st)
{
    hipLaunchKernelGGL(k_epoch_end, dim3(std::max(1, std::min(1024, (a.n_jobs + 255) / 256))), dim3(256), 0, st, a,
                       (int32_t*)a.active, (int32_t*)a.n_active);
    hipLaunchKernelGGL(k_epoch_advance, dim3(1), dim3(64), 0, st, (int32_t*)a.epoch, a.n_active, remaining);
    return hipGetLastError();
}

size_t solve_state_lds_bytes(int32_t max_n) { return Layout::state_bytes(max_n); }

hipError_t launch_solve(const SolveArgs& a, int clause_mem, int waves, size_t lds_bytes, hipStream_t st)
{
    // waves is a multiple of the wavefronts per workgroup (solve_waves_per_block)
    const int wpb = solve_waves_per_block(lds_bytes);
    if (waves % wpb != 0 || (size_t)a.lds_per_wave != lds_bytes) return hipErrorInvalidValue;
    const dim3 grid(waves / wpb), block(kWave * wpb);
    if (clause_mem == 1)
        hipLaunchKernelGGL(k_solve<1>, grid, block, lds_bytes * wpb, st, a);
    else if (clause_mem == 3)
        hipLaunchKernelGGL(k_solve<3>, grid, block, lds_bytes * wpb, st, a);
    else if (clause_mem == 4)
        hipLaunchKernelGGL(k_solve<4>, grid, block, 0, st, a);
    else
        hipLaunchKernelGGL(k_solve<2>, grid, block, lds_bytes * wpb, st, a);
    return hipGetLastError();
}

int solve_waves_per_block(size_t lds_bytes)
{
    int wpb = kSolveWavesPerBlock;
    while (wpb > 1 && (size_t)wpb * lds_bytes > 64 * 1024) wpb >>= 1;
    return wpb;
}

}  // namespace gpsat

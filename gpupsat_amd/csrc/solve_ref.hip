// solve_ref.hip -- the reference's own sequential search, on the device, one instance per
// lane: the trajectory mode.
//
// This follows run_sequential / SATSolver::solve (reference src/SATSolver/SATSolver.cu:67-218)
// step for step, so that verdicts, iteration counts and the UNDEF set under MAX_ITERATIONS
// are the reference scheme's, not just verdicts:
//   * decisions: the last variable of the free-variable list, positive polarity
//     (DecisionMaker.cu:45-86); the list starts in ascending order, freed variables are
//     appended (VariablesStateHandler.cu:19-30, 218-248, 351), removal keeps the order;
//   * propagation: two watched literals per clause, watch lists walked in insertion order
//     (WatchedClausesList.cu:46-309);
//   * learning: the decision scheme -- a depth-first walk back from the conflict through
//     reasons of the current level; lower-level literals and current-level roots form the
//     clause (GraphAnalyzer.cu:31-138); backjump to the second-highest level and assert the
//     current-level literal; a unit clause backtracks to 0 and is asserted there
//     (ConflictAnalyzer.cu:105-151);
//   * learned clauses of 2 <= size < max(largest clause, 100) are kept in a 20-slot
//     round-robin database (LearntClauseRepository.cu:15-33, Configs.cuh:94);
//   * restarts: geometric, 100 conflicts then x1.3 (GeometricRestartsManager.cu:3-31);
//   * MAX_ITERATIONS: after that many decisions the answer is UNDEF (Configs.cuh:23).
// Lanes run independent instances (the reference's "thread per job"); every per-instance
// array lives in an HBM slice of the lane's own.  Capacities are fixed per call; running
// out of one is reported as status 4 and never written past.
#include "../../include/gpsat.h"
#include "common.hpp"
#include "gpsat_internal.hpp"
#include "solver.hpp"

namespace gpsat {

namespace {

constexpr int32_t kRefDb = 20;             // MAX_LEARNT_CLAUSES_PER_THREAD (Configs.cuh:94)
constexpr int32_t kRefMinImpl = 100;       // MIN_IMPLICATION_PER_VAR (main.cu:148)
constexpr int32_t kRefSlack = 64;          // watch slots per literal beyond its occurrences

// One lane's solver state; every array is a slice of the lane's workspace.
struct RS {
    int32_t n, m;
    int32_t *val, *level, *reason, *isdec, *trail, *fprev, *fnext, *infree, *seen, *learnt, *stack;
    int32_t *alits, *cstart, *clen, *active, *w0, *w1;
    int32_t *wstart, *wcap, *wn, *wdata;
    int32_t nl, nc, ccap, lcap;
    int32_t ntrail, qhead, dl;
    int32_t fhead, ftail, nfree;
    bool overflow;
};

__device__ __forceinline__ int32_t lit_val(const RS& s, int32_t l)
{
    const int32_t x = s.val[l >> 1];
    return x < 0 ? -1 : ((l & 1) ? x : 1 - x);
}

__device__ void fl_append(RS& s, int32_t v)
{
    if (s.infree[v]) return;
    s.infree[v] = 1;
    s.nfree++;
    s.fprev[v] = s.ftail;
    s.fnext[v] = -1;
    if (s.ftail >= 0) s.fnext[s.ftail] = v;
    else s.fhead = v;
    s.ftail = v;
}

__device__ void fl_remove(RS& s, int32_t v)
{
    if (!s.infree[v]) return;
    s.infree[v] = 0;
    s.nfree--;
    if (s.fprev[v] >= 0) s.fnext[s.fprev[v]] = s.fnext[v];
    else s.fhead = s.fnext[v];
    if (s.fnext[v] >= 0) s.fprev[s.fnext[v]] = s.fprev[v];
    else s.ftail = s.fprev[v];
}

__device__ void rs_assign(RS& s, int32_t l, int32_t lvl, int32_t reason, int32_t decision)
{
    const int32_t v = l >> 1;
    s.val[v] = l & 1;
    s.level[v] = lvl;
    s.reason[v] = reason;
    s.isdec[v] = decision;
    s.trail[s.ntrail++] = l;
    fl_remove(s, v);
}

// Append clause c to the watch list of l; a full list first drops its stale entries
// (clauses no longer watching l, or deactivated), which the walk would skip anyway, so
// the order of the live entries -- and the trajectory -- is unchanged.
__device__ void w_push(RS& s, int32_t l, int32_t c)
{
    int32_t* W = s.wdata + s.wstart[l];
    if (s.wn[l] == s.wcap[l]) {
        int32_t j = 0;
        for (int32_t i = 0; i < s.wn[l]; i++) {
            const int32_t d = W[i];
            if (s.active[d] && (s.w0[d] == l || s.w1[d] == l)) W[j++] = d;
        }
        s.wn[l] = j;
        if (j == s.wcap[l]) {
            s.overflow = true;
            return;
        }
    }
    W[s.wn[l]++] = c;
}

__device__ void rs_watch(RS& s, int32_t c)
{
    const int32_t* cl = s.alits + s.cstart[c];
    s.w0[c] = cl[0];
    s.w1[c] = cl[1];
    w_push(s, cl[0], c);
    w_push(s, cl[1], c);
}

// Watched-literal propagation to a fixpoint; the conflicting clause or -1.
__device__ int32_t rs_propagate(RS& s)
{
    while (s.qhead < s.ntrail && !s.overflow) {
        const int32_t p = s.trail[s.qhead++];
        const int32_t fl = p ^ 1;
        int32_t* W = s.wdata + s.wstart[fl];
        const int32_t cnt = s.wn[fl];
        int32_t i = 0, j = 0, confl = -1;
        while (i < cnt) {
            const int32_t c = W[i++];
            if (!s.active[c] || (s.w0[c] != fl && s.w1[c] != fl)) continue;   // stale
            const int32_t other = (s.w0[c] == fl) ? s.w1[c] : s.w0[c];
            if (confl >= 0 || lit_val(s, other) == 1) {
                W[j++] = c;
                continue;
            }
            const int32_t* cl = s.alits + s.cstart[c];
            const int32_t len = s.clen[c];
            bool moved = false;
            for (int32_t k = 0; k < len; k++) {
                const int32_t q = cl[k];
                if (q == s.w0[c] || q == s.w1[c]) continue;
                if (lit_val(s, q) != 0) {
                    if (s.w0[c] == fl) s.w0[c] = q;
                    else s.w1[c] = q;
                    w_push(s, q, c);
                    moved = true;
                    break;
                }
            }
            if (moved) continue;
            W[j++] = c;
            if (lit_val(s, other) == 0) confl = c;
            else rs_assign(s, other, s.dl, c, 0);
        }
        s.wn[fl] = j;
        if (confl >= 0) return confl;
    }
    return -1;
}

// VariablesStateHandler::backtrack_to (VariablesStateHandler.cu:367-399): implications
// above the level return to the free list first, in trail order, then decisions.
__device__ void rs_backtrack(RS& s, int32_t lvl)
{
    for (int32_t i = 0; i < s.ntrail; i++) {
        const int32_t v = s.trail[i] >> 1;
        if (s.level[v] > lvl && !s.isdec[v]) fl_append(s, v);
    }
    for (int32_t i = 0; i < s.ntrail; i++) {
        const int32_t v = s.trail[i] >> 1;
        if (s.level[v] > lvl && s.isdec[v]) fl_append(s, v);
    }
    int32_t keep = 0;
    for (int32_t i = 0; i < s.ntrail; i++) {
        const int32_t l = s.trail[i], v = l >> 1;
        if (s.level[v] > lvl) {
            s.val[v] = -1;
            s.reason[v] = -1;
        } else {
            s.trail[keep++] = l;
        }
    }
    s.ntrail = keep;
    s.qhead = keep;
    s.dl = lvl;
}

__device__ int32_t arena_add(RS& s, const int32_t* l, int32_t n)
{
    if (s.nl + n > s.lcap || s.nc + 1 > s.ccap) {
        s.overflow = true;
        return -1;
    }
    for (int32_t k = 0; k < n; k++) s.alits[s.nl + k] = l[k];
    s.cstart[s.nc] = s.nl;
    s.clen[s.nc] = n;
    s.active[s.nc] = 0;
    s.nl += n;
    return s.nc++;
}

}  // namespace

__global__ __launch_bounds__(64) void k_solve_ref(RefArgs a)
{
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= a.n_inst) return;
    const InstMeta mi = a.meta[inst];
    const RefMeta rm = a.rmeta[inst];
    const int32_t n = mi.n, m = mi.m;
    const int32_t* lits = a.lits + mi.lit_base;
    const int32_t* coff = a.coff + mi.coff_base;
    RS s;
    s.n = n;
    s.m = m;
    int32_t* w = a.ws + rm.ws;
    auto take = [&](int64_t k) { int32_t* p = w; w += k; return p; };
    const int32_t nv = n > 0 ? n : 1;
    s.val = take(nv); s.level = take(nv); s.reason = take(nv); s.isdec = take(nv); s.trail = take(nv);
    s.fprev = take(nv); s.fnext = take(nv); s.infree = take(nv); s.seen = take(nv);
    s.learnt = take(nv + 1); s.stack = take(nv + 1);
    s.ccap = rm.ccap;
    s.lcap = rm.lcap;
    s.alits = take(rm.lcap); s.cstart = take(rm.ccap); s.clen = take(rm.ccap); s.active = take(rm.ccap);
    s.w0 = take(rm.ccap); s.w1 = take(rm.ccap);
    s.wstart = take(2 * nv); s.wcap = take(2 * nv); s.wn = take(2 * nv); s.wdata = take(rm.wtot);
    s.overflow = false;

    // watch-list capacity per literal: its occurrences plus slack
    for (int32_t l = 0; l < 2 * nv; l++) s.wcap[l] = kRefSlack;
    for (int32_t k = 0; k < coff[m]; k++) s.wcap[lits[k]]++;
    int32_t pos = 0;
    for (int32_t l = 0; l < 2 * nv; l++) {
        s.wstart[l] = pos;
        s.wn[l] = 0;
        pos += s.wcap[l];
    }
    s.ntrail = s.qhead = s.dl = 0;
    s.fhead = s.ftail = -1;
    s.nfree = 0;
    const uint8_t* dead = a.dead ? a.dead + a.var_base[inst] : nullptr;
    for (int32_t v = 0; v < n; v++) {
        s.val[v] = -1;
        s.level[v] = 0;
        s.reason[v] = -1;
        s.isdec[v] = 0;
        s.infree[v] = 0;
        if (!(dead && dead[v])) fl_append(s, v);
    }
    s.nl = coff[m];
    s.nc = m;
    for (int32_t k = 0; k < s.nl; k++) s.alits[k] = lits[k];
    int32_t status = GPSAT_UNDEF;
    int32_t largest = 0;
    for (int32_t c = 0; c < m; c++) {
        const int32_t len = coff[c + 1] - coff[c];
        s.cstart[c] = coff[c];
        s.clen[c] = len;
        s.active[c] = 1;
        if (len > largest) largest = len;
        if (len == 0) {
            status = GPSAT_UNSAT;
        } else if (len == 1) {
            const int32_t l = lits[coff[c]];
            const int32_t lv = lit_val(s, l);
            if (lv == 0) status = GPSAT_UNSAT;
            else if (lv < 0) rs_assign(s, l, 0, -1, 0);
        } else {
            rs_watch(s, c);
        }
    }
    const int32_t max_impl = largest > kRefMinImpl ? largest : kRefMinImpl;
    int32_t db[kRefDb];
    int32_t db_n = 0, db_pos = 0;
    int32_t restart_thr = 100, restart_cnt = 0;
    const float restart_factor = 1.3f;
    int64_t st_dec = 0, st_conf = 0, st_learn = 0, st_iter = 0, st_rest = 0;
    if (status == GPSAT_UNDEF && rs_propagate(s) >= 0) status = GPSAT_UNSAT;
    if (status == GPSAT_UNDEF && s.nfree == 0) status = GPSAT_SAT;
    int32_t iterations = 0;
    while (status == GPSAT_UNDEF && !s.overflow) {
        // DecisionMaker::decide: the last free var, positive, one level up
        const int32_t v = s.ftail;
        s.dl++;
        st_dec++;
        rs_assign(s, 2 * v + 1, s.dl, -1, 1);
        int32_t n_conf = 0, confl;
        while ((confl = rs_propagate(s)) >= 0) {
            n_conf++;
            st_conf++;
            if (s.dl <= 0) {
                status = GPSAT_UNSAT;
                break;
            }
            // decision-scheme analysis: depth first through current-level reasons
            int32_t nl = 0, sp = 0, hi_lit = -1, second = -2;
            const int32_t cur = s.dl;
            for (int32_t u = 0; u < n; u++) s.seen[u] = 0;
            s.stack[sp++] = confl;
            while (sp > 0) {
                const int32_t c = s.stack[--sp];
                const int32_t* cl = s.alits + s.cstart[c];
                const int32_t len = s.clen[c];
                for (int32_t k = 0; k < len; k++) {
                    const int32_t u = cl[k] >> 1;
                    if (s.seen[u]) continue;
                    s.seen[u] = 1;
                    const int32_t assigned = 2 * u + s.val[u];
                    if (s.level[u] < cur || s.reason[u] < 0) {
                        s.learnt[nl++] = assigned ^ 1;
                        if (s.level[u] >= cur) hi_lit = assigned ^ 1;
                        else if (s.level[u] > second) second = s.level[u];
                    } else {
                        s.stack[sp++] = s.reason[u];
                    }
                }
            }
            st_learn++;
            if (nl == 1) {
                rs_backtrack(s, 0);
                rs_assign(s, s.learnt[0], 0, -1, 0);
            } else {
                if (second < 0) second = 0;
                const int32_t cid = arena_add(s, s.learnt, nl);
                if (cid < 0) break;
                rs_backtrack(s, second);
                rs_assign(s, hi_lit, s.dl, cid, 0);
                if (nl < max_impl) {
                    // watch the asserted literal and a literal of the backjump level
                    int32_t* cl = s.alits + s.cstart[cid];
                    for (int32_t k = 0; k < nl; k++)
                        if (cl[k] == hi_lit) {
                            const int32_t t = cl[0];
                            cl[0] = cl[k];
                            cl[k] = t;
                        }
                    int32_t bi = 0, best = -1;
                    for (int32_t k = 1; k < nl; k++) {
                        const int32_t lv = s.level[cl[k] >> 1];
                        if (lv > best) {
                            best = lv;
                            bi = k;
                        }
                    }
                    const int32_t t = cl[1];
                    cl[1] = cl[bi];
                    cl[bi] = t;
                    if (db_n < kRefDb) {
                        db[db_n++] = cid;
                    } else {
                        s.active[db[db_pos]] = 0;
                        db[db_pos] = cid;
                        if (++db_pos >= db_n) db_pos = 0;
                    }
                    s.active[cid] = 1;
                    rs_watch(s, cid);
                }
            }
            if (s.overflow) break;
        }
        if (status != GPSAT_UNDEF || s.overflow) break;
        if (s.nfree == 0) {
            status = GPSAT_SAT;
            break;
        }
        restart_cnt += n_conf;
        if (restart_cnt >= restart_thr) {
            restart_cnt = 0;
            restart_thr = (int32_t)((float)restart_thr * restart_factor);
            rs_backtrack(s, 0);
            st_rest++;
        }
        iterations++;
        st_iter = iterations;
        if (a.max_iterations > 0 && iterations > a.max_iterations) break;
    }
    if (s.overflow) status = 4;
    a.status[inst] = status;
    int64_t* so = a.stats + (int64_t)inst * 5;
    so[0] = st_dec;
    so[1] = st_conf;
    so[2] = st_learn;
    so[3] = st_iter;
    so[4] = st_rest;
    uint32_t* row = a.model_rows + mi.word_base;
    const int32_t W = words_per_row(n);
    for (int32_t wd = 0; wd < W; wd++) {
        uint32_t x = 0;
        if (status == GPSAT_SAT)
            for (int32_t b = 0; b < 32; b++) {
                const int32_t l = 32 * wd + b;
                if (l < 2 * n && lit_val(s, l) == 1) x |= 1u << b;
            }
        row[wd] = x;
    }
}

// Per-instance workspace ints for the trajectory mode.
int64_t solve_ref_ws_ints(int32_t n, int32_t m, int32_t L, int32_t ccap, int32_t lcap)
{
    const int64_t nv = n > 0 ? n : 1;
    const int64_t wtot = (int64_t)2 * nv * kRefSlack + L;
    return 9 * nv + 2 * (nv + 1) + lcap + 5 * (int64_t)ccap + 6 * nv + wtot;
}

int32_t solve_ref_wtot(int32_t n, int32_t L) { return 2 * (n > 0 ? n : 1) * kRefSlack + L; }

hipError_t launch_solve_ref(const RefArgs& a, hipStream_t st)
{
    if (a.n_inst == 0) return hipSuccess;
    hipLaunchKernelGGL(k_solve_ref, dim3((a.n_inst + 63) / 64), dim3(64), 0, st, a);
    return hipGetLastError();
}

}  // namespace gpsat

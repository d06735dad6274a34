// solve_ref.hip -- the reference's own sequential search, on the device, one instance per
// lane: the trajectory mode.
//
// Step for step the search of run_sequential / SATSolver::solve (reference
// src/SATSolver/SATSolver.cu:67-218) with its default strategies, over the same structures:
//   * the free-variable list in order: ascending at the start, freed variables appended,
//     removal keeping the order; decisions take its last variable, positive
//     (VariablesStateHandler.cu:4-33, 325-399; DecisionMaker.cu:45-86);
//   * two watched literals kept in per-variable lists, newest first: a clause watches the two
//     literals whose variables have the fewest watched clauses; unit, satisfied-by-implication
//     and conflicting clauses move to a blocked list, re-watched on backtrack when unresolved
//     (ClauseListStructure.cu:39-243);
//   * propagation of a decision: the decided variable's list as it stands, clause by clause,
//     against the assignment plus the implications found so far, then depth first into each
//     implication this step found; all found implications are applied afterwards, also when a
//     conflict ended the walk (WatchedClausesList.cu:46-309,
//     ConflictAnalyzerWithWatchedLits.cu:113-139);
//   * the implication graph: one vertex per variable and a conflict vertex, backward edges in
//     insertion order; removing a vertex drops the first edge to it from every other vertex
//     (CUDAListGraph.cu, GraphStructure.cu:79-330);
//   * decision-scheme learning: depth first from the conflict vertex, a neighbour below the
//     level of the vertex being expanded joins the clause, others are expanded
//     (GraphAnalyzer.cu:31-138); backjump through Backtracker::handle_backtrack
//     (Backtracker.cu:37-75) and the asserting implication (ConflictAnalyzer.cu:34-171);
//   * learned clauses of 2 <= size < max(largest clause, 100) go to a 20-slot round-robin
//     repository, replacing the oldest's watch record (LearntClausesManager.cu:16-41,
//     LearntClauseRepository.cu:15-33, WatchedClausesList.cu:33-46);
//   * geometric restarts 100 x 1.3 (GeometricRestartsManager.cu:3-31) and MAX_ITERATIONS.
// The CPU restatement of the same steps is oracle/refsolve.c; the two give the same status,
// model and counters instance by instance.  Lanes run independent instances (the reference's
// thread per job); every array lives in an HBM slice of the lane's own, sized from the
// instance, and the recursions of the reference run on explicit stacks in that slice.  A
// vertex with more backward edges than the reference's per-vertex capacity
// (max implication per var + 1) is reported as status 4, as the oracle reports it.
#include "../../include/gpsat.h"
#include "common.hpp"
#include "gpsat_internal.hpp"
#include "solver.hpp"

namespace gpsat {

namespace {

constexpr int32_t kSlots = 20;             // MAX_LEARNT_CLAUSES_PER_THREAD (Configs.cuh:94)
constexpr int32_t kNullLevel = -234;       // NULL_DECISION_LEVEL (Configs.cuh:24)
constexpr int32_t kIntMax = 0x7fffffff;

struct Dec {
    int32_t lit, level, branched, formula;
};

// One lane's solver state; every array is a slice of the lane's workspace.
struct RS {
    int32_t n, m, nw, ecap, max_impl;
    const int32_t* olits;      // original literals of the instance
    const int32_t* ocoff;      // original clause offsets (relative)
    // watched-clause records (slots): clause id, watched indices, literal start (>= 0: original
    // literal offset; < 0: -1 - offset into lstore) and length
    int32_t *wc_clause, *wc_w1, *wc_w2, *wc_start, *wc_len;
    int32_t n_wc;
    int32_t* lstore;           // 21 learned-clause slots of n literals (20 repository + current)
    int32_t *lrep_id, *lrep_wc; int32_t nlrep, lpos;
    int32_t *nd_val, *nd_next; int32_t nd_free;
    int32_t* vlist;            // per var: head node
    int32_t blocked;
    int32_t* st;               // per var: 1 true, 0 false, -1 undef
    int32_t* freev; int32_t nfree;
    int32_t* decs; int32_t ndec;      // 4 ints per decision: lit, level, branched, formula
    int32_t* imps; int32_t nimp;
    int32_t level;
    int32_t *vx, *vx_nb, *vx_set, *vx_flag, *edges, *vidx;   // vx: 4 ints per vertex
    int32_t nvidx;
    int32_t* pend; int32_t npend;     // 4 ints: lit, level, formula, watch slot
    int32_t* pmark;                   // per literal
    int32_t* frames;                  // new_decision stack: lit, formula, start, found, i, entered
    int32_t* snap;
    int32_t *lbuf, *astack;
    int32_t next_id;
    bool overflow;
};

__host__ __device__ __forceinline__ const int32_t* wc_lits(const RS& s, int32_t w)
{
    const int32_t st = s.wc_start[w];
    return st >= 0 ? s.olits + st : s.lstore + (-1 - st);
}

__host__ __device__ __forceinline__ Dec get_dec(const int32_t* p)
{
    return Dec{p[0], p[1], p[2], p[3]};
}
__host__ __device__ __forceinline__ void put_dec(int32_t* p, const Dec& d)
{
    p[0] = d.lit;
    p[1] = d.level;
    p[2] = d.branched;
    p[3] = d.formula;
}

// ---- node pool and lists (newest first)
__host__ __device__ int32_t node_new(RS& s, int32_t val)
{
    const int32_t k = s.nd_free;   // the pool holds 3 nodes per watch slot: never exhausted
    s.nd_free = s.nd_next[k];
    s.nd_val[k] = val;
    return k;
}
__host__ __device__ __forceinline__ void node_free(RS& s, int32_t k)
{
    s.nd_next[k] = s.nd_free;
    s.nd_free = k;
}
__host__ __device__ void add_first(RS& s, int32_t* head, int32_t val)
{
    const int32_t k = node_new(s, val);
    s.nd_next[k] = *head;
    *head = k;
}
__host__ __device__ int32_t list_size(const RS& s, int32_t head)
{
    int32_t c = 0;
    for (int32_t k = head; k >= 0; k = s.nd_next[k]) c++;
    return c;
}
__host__ __device__ void remove_clause(RS& s, int32_t* head, int32_t c)
{
    int32_t prev = -1;
    for (int32_t k = *head; k >= 0; prev = k, k = s.nd_next[k])
        if (s.wc_clause[s.nd_val[k]] == c) {
            if (prev < 0) *head = s.nd_next[k];
            else s.nd_next[prev] = s.nd_next[k];
            node_free(s, k);
            return;
        }
}

// ---- variable state
__host__ __device__ __forceinline__ int32_t lit_status(const RS& s, int32_t l)
{
    const int32_t x = s.st[l >> 1];
    if (x < 0) return GPSAT_UNDEF;
    return (l & 1) == x ? GPSAT_SAT : GPSAT_UNSAT;
}
__host__ __device__ __forceinline__ bool consistent(const RS& s, int32_t l) { return lit_status(s, l) != GPSAT_UNSAT; }
__host__ __device__ void free_remove(RS& s, int32_t v)
{
    for (int32_t i = 0; i < s.nfree; i++)
        if (s.freev[i] == v) {
            for (int32_t j = i; j + 1 < s.nfree; j++) s.freev[j] = s.freev[j + 1];
            s.nfree--;
            return;
        }
}
__host__ __device__ __forceinline__ void free_add(RS& s, int32_t v)
{
    if (s.nfree < s.n) s.freev[s.nfree++] = v;
}
__host__ __device__ void new_implication(RS& s, const Dec& d)
{
    put_dec(s.imps + 4 * s.nimp++, d);
    free_remove(s, d.lit >> 1);
    s.st[d.lit >> 1] = d.lit & 1;
}
__host__ __device__ void new_decision(RS& s, const Dec& d)
{
    put_dec(s.decs + 4 * s.ndec++, d);
    free_remove(s, d.lit >> 1);
    s.st[d.lit >> 1] = d.lit & 1;
}
__host__ __device__ bool formula_implied(const RS& s, int32_t v)
{
    for (int32_t i = 0; i < s.nimp; i++)
        if ((s.imps[4 * i] >> 1) == v) return s.imps[4 * i + 3] != 0;
    return false;
}
__host__ __device__ bool should_implicate_from_formula(const RS& s, int32_t w, int32_t unit)
{
    const int32_t* cl = wc_lits(s, w);
    int32_t count = 0;
    for (int32_t i = 0; i < s.wc_len[w]; i++)
        if ((cl[i] >> 1) == (unit >> 1) || formula_implied(s, cl[i] >> 1)) count++;
    return count == s.wc_len[w];
}
__host__ __device__ bool is_unresolved(const RS& s, int32_t w)
{
    const int32_t* cl = wc_lits(s, w);
    const int32_t n = s.wc_len[w];
    if (n == 1) return false;
    int32_t n_unsat = 0;
    for (int32_t i = 0; i < n; i++) {
        const int32_t ls = lit_status(s, cl[i]);
        if (ls == GPSAT_SAT) return false;
        if (ls == GPSAT_UNSAT) n_unsat++;
    }
    return n_unsat < n - 1;
}

// ---- watch structure
__host__ __device__ void watch_clause(RS& s, int32_t w, bool test)
{
    const int32_t* cl = wc_lits(s, w);
    int32_t sv = kIntMax - 1, ssv = kIntMax, sm = 0, ssm = 0;
    for (int32_t i = 0; i < s.wc_len[w]; i++) {
        if (test && !consistent(s, cl[i])) continue;
        const int32_t size = list_size(s, s.vlist[cl[i] >> 1]);
        if (size < sv) {
            ssm = sm;
            ssv = sv;
            sv = size;
            sm = i;
        } else if (size < ssv) {
            ssv = size;
            ssm = i;
        }
    }
    s.wc_w1[w] = sm;
    s.wc_w2[w] = ssm;
    add_first(s, &s.vlist[cl[sm] >> 1], w);
    add_first(s, &s.vlist[cl[ssm] >> 1], w);
}
__host__ __device__ void block_clause(RS& s, int32_t w)
{
    const int32_t* cl = wc_lits(s, w);
    const int32_t c = s.wc_clause[w];
    if (s.wc_w1[w] >= 0) remove_clause(s, &s.vlist[cl[s.wc_w1[w]] >> 1], c);
    if (s.wc_w2[w] >= 0) remove_clause(s, &s.vlist[cl[s.wc_w2[w]] >> 1], c);
    add_first(s, &s.blocked, w);
}
__host__ __device__ void switch_watch(RS& s, int32_t old_i, int32_t new_i, int32_t w)
{
    const int32_t* cl = wc_lits(s, w);
    if (s.wc_w1[w] == old_i) s.wc_w1[w] = new_i;
    else s.wc_w2[w] = new_i;
    remove_clause(s, &s.vlist[cl[old_i] >> 1], s.wc_clause[w]);
    add_first(s, &s.vlist[cl[new_i] >> 1], w);
}
__host__ __device__ void unblock_all(RS& s)
{
    int32_t prev = -1;
    for (int32_t k = s.blocked; k >= 0;) {
        const int32_t nxt = s.nd_next[k], w = s.nd_val[k];
        if (is_unresolved(s, w)) {
            watch_clause(s, w, true);
            if (prev < 0) s.blocked = nxt;
            else s.nd_next[prev] = nxt;
            node_free(s, k);
        } else {
            prev = k;
        }
        k = nxt;
    }
}
__host__ __device__ void new_clause_checked(RS& s, int32_t w)
{
    const int32_t* cl = wc_lits(s, w);
    const int32_t n = s.wc_len[w];
    int32_t n_unsat = 0, unit = -1;
    bool sat = false;
    for (int32_t i = 0; i < n && !sat; i++) {
        const int32_t ls = lit_status(s, cl[i]);
        if (ls == GPSAT_SAT) sat = true;
        else if (ls == GPSAT_UNSAT) n_unsat++;
        else unit = cl[i];
    }
    if (sat || n_unsat != n - 1) unit = -1;
    if (!sat && (n_unsat == n || unit >= 0)) {
        s.wc_w1[w] = s.wc_w2[w] = -1;   // in no var list yet: blocking only adds it to the blocked list
        block_clause(s, w);
    } else {
        watch_clause(s, w, false);
    }
}
// purge_clause for the clause in watch slot w (ClauseListStructure.cu:191-243)
__host__ __device__ void purge(RS& s, int32_t w)
{
    const int32_t c = s.wc_clause[w];
    int32_t prev = -1;
    for (int32_t k = s.blocked; k >= 0; prev = k, k = s.nd_next[k])
        if (s.wc_clause[s.nd_val[k]] == c) {
            if (prev < 0) s.blocked = s.nd_next[k];
            else s.nd_next[prev] = s.nd_next[k];
            node_free(s, k);
            return;
        }
    const int32_t* cl = wc_lits(s, w);
    int32_t other = -1;
    for (int32_t i = 0; i < s.wc_len[w] && other < 0; i++) {
        const int32_t v = cl[i] >> 1;
        int32_t pv = -1;
        for (int32_t k = s.vlist[v]; k >= 0; pv = k, k = s.nd_next[k]) {
            const int32_t x = s.nd_val[k];
            if (s.wc_clause[x] == c) {
                const int32_t v1 = cl[s.wc_w1[x]] >> 1, v2 = cl[s.wc_w2[x]] >> 1;
                other = v1 == v ? v2 : v1;
                if (pv < 0) s.vlist[v] = s.nd_next[k];
                else s.nd_next[pv] = s.nd_next[k];
                node_free(s, k);
                break;
            }
        }
    }
    if (other >= 0) remove_clause(s, &s.vlist[other], c);
}

// ---- propagation over the pending implications
__host__ __device__ __forceinline__ bool cur_consistent(const RS& s, int32_t l) { return consistent(s, l) && !s.pmark[l ^ 1]; }
__host__ __device__ int32_t pending_status(const RS& s, int32_t l)
{
    const int32_t ls = lit_status(s, l);
    if (ls != GPSAT_UNDEF) return ls;
    if (s.pmark[l]) return GPSAT_SAT;
    if (s.pmark[l ^ 1]) return GPSAT_UNSAT;
    return GPSAT_UNDEF;
}
__host__ __device__ void handle_implication(RS& s, int32_t w, int32_t unit, int32_t from_formula)
{
    if (pending_status(s, unit) == GPSAT_UNDEF) {
        int32_t* p = s.pend + 4 * s.npend++;
        p[0] = unit;
        p[1] = s.level;
        p[2] = from_formula ? (should_implicate_from_formula(s, w, unit) ? 1 : 0) : 0;
        p[3] = w;
        s.pmark[unit] = 1;
    }
    block_clause(s, w);
}
__host__ __device__ bool process_clause(RS& s, int32_t w, int32_t from_formula)
{
    const int32_t* cl = wc_lits(s, w);
    const int32_t n = s.wc_len[w];
    const int32_t l1 = cl[s.wc_w1[w]], l2 = cl[s.wc_w2[w]];
    const bool c1 = cur_consistent(s, l1), c2 = cur_consistent(s, l2);
    if (c1 && c2) return false;
    int32_t pending = (!c1 && !c2) ? 2 : 1, o1 = -1, o2 = -1;
    for (int32_t i = 0; pending > 0 && i < n; i++) {
        const int32_t l = cl[i];
        if (l != l1 && l != l2 && cur_consistent(s, l)) {
            if (o1 == -1) o1 = i;
            else o2 = i;
            pending--;
        }
    }
    if (c1 || c2) {
        if (o1 == -1) handle_implication(s, w, cl[c1 ? s.wc_w1[w] : s.wc_w2[w]], from_formula);
        else switch_watch(s, !c1 ? s.wc_w1[w] : s.wc_w2[w], o1, w);
        return false;
    }
    if (o1 != -1 && o2 != -1) {
        switch_watch(s, s.wc_w1[w], o1, w);
        switch_watch(s, s.wc_w2[w], o2, w);
        return false;
    }
    if (o1 != -1 || o2 != -1) {
        handle_implication(s, w, cl[o1 != -1 ? o1 : o2], from_formula);
        return false;
    }
    block_clause(s, w);
    return true;
}
// WatchedClausesList::new_decision as an explicit stack: a frame walks the snapshot of its
// literal's var list, then descends into the implications it found, in order.  Returns the
// watch slot of the conflicting clause or -1.
__host__ __device__ int32_t wl_new_decision(RS& s, int32_t lit, int32_t from_formula)
{
    int32_t top = 0;
    int32_t* f = s.frames;
    f[0] = lit;
    f[1] = from_formula;
    f[5] = 0;
    for (;;) {
        int32_t* F = f + 6 * top;
        if (!F[5]) {
            F[5] = 1;
            int32_t cnt = 0;
            for (int32_t k = s.vlist[F[0] >> 1]; k >= 0; k = s.nd_next[k]) s.snap[cnt++] = s.nd_val[k];
            F[2] = s.npend;
            for (int32_t i = 0; i < cnt; i++)
                if (process_clause(s, s.snap[i], F[1])) return s.snap[i];
            F[3] = s.npend - F[2];
            F[4] = 0;
        }
        if (F[4] < F[3]) {
            const int32_t* p = s.pend + 4 * (F[2] + F[4]);
            F[4]++;
            top++;
            int32_t* G = f + 6 * top;
            G[0] = p[0];
            G[1] = p[2];
            G[5] = 0;
        } else {
            if (top == 0) return -1;
            top--;
        }
    }
}

// ---- implication graph
__host__ __device__ __forceinline__ void vidx_add(RS& s, int32_t v)
{
    if (s.nvidx < s.n + 1) s.vidx[s.nvidx++] = v;
}
__host__ __device__ void graph_set(RS& s, const Dec& d)
{
    const int32_t v = d.lit >> 1;
    put_dec(s.vx + 4 * v, d);
    s.vx_nb[v] = 0;
    s.vx_set[v] = 1;
    vidx_add(s, v);
}
__host__ __device__ __forceinline__ void graph_edge(RS& s, int32_t src, int32_t dst)
{
    if (s.vx_nb[dst] >= s.ecap) {
        s.overflow = true;
        return;
    }
    s.edges[(int64_t)dst * s.ecap + s.vx_nb[dst]++] = src;
}
__host__ __device__ void graph_remove_vertex(RS& s, int32_t v)
{
    s.vx_set[v] = 0;
    s.vx_flag[v] = 0;
    s.vx_nb[v] = 0;
    for (int32_t i = 0; i < s.nvidx; i++)
        if (s.vidx[i] == v) {
            for (int32_t j = i; j + 1 < s.nvidx; j++) s.vidx[j] = s.vidx[j + 1];
            s.nvidx--;
            break;
        }
    for (int32_t i = 0; i < s.nvidx; i++) {
        const int32_t o = s.vidx[i];
        int32_t* e = s.edges + (int64_t)o * s.ecap;
        for (int32_t j = 0; j < s.vx_nb[o]; j++)
            if (e[j] == v) {
                for (int32_t q = j; q + 1 < s.vx_nb[o]; q++) e[q] = e[q + 1];
                s.vx_nb[o]--;
                break;
            }
    }
}
__host__ __device__ void graph_backtrack(RS& s, int32_t lvl)
{
    for (int32_t i = 0; i < s.nvidx; i++) {
        const int32_t v = s.vidx[i];
        if (s.vx[4 * v + 1] > lvl && !s.vx[4 * v + 3]) {
            graph_remove_vertex(s, v);
            i--;
        }
    }
}
__host__ __device__ void add_implication(RS& s, const Dec& d, const int32_t* cl, int32_t len)
{
    graph_set(s, d);
    for (int32_t i = 0; i < len; i++)
        if (cl[i] != d.lit) graph_edge(s, cl[i] >> 1, d.lit >> 1);
    new_implication(s, d);
}

// ---- backtracking
__host__ __device__ void backtrack_to(RS& s, int32_t lvl)
{
    s.level = lvl;
    for (int32_t i = 0; i < s.nimp; i++)
        if (s.imps[4 * i + 1] > lvl) {
            const int32_t v = s.imps[4 * i] >> 1;
            for (int32_t j = i; j + 1 < s.nimp; j++)
                for (int32_t q = 0; q < 4; q++) s.imps[4 * j + q] = s.imps[4 * (j + 1) + q];
            s.nimp--;
            free_add(s, v);
            s.st[v] = -1;
            i--;
        }
    for (int32_t i = 0; i < s.ndec; i++)
        if (s.decs[4 * i + 1] > lvl) {
            const int32_t v = s.decs[4 * i] >> 1;
            for (int32_t j = i; j + 1 < s.ndec; j++)
                for (int32_t q = 0; q < 4; q++) s.decs[4 * j + q] = s.decs[4 * (j + 1) + q];
            s.ndec--;
            free_add(s, v);
            s.st[v] = -1;
            i--;
        }
    graph_backtrack(s, lvl);
    unblock_all(s);
}
__host__ __device__ int32_t handle_backtrack(RS& s, int32_t lvl, bool switch_literal)
{
    Dec tb{0, kNullLevel, 0, 0};
    for (int32_t i = s.ndec - 1; i >= 0; i--)
        if (!s.decs[4 * i + 2] && s.decs[4 * i + 1] <= lvl) {
            tb = get_dec(s.decs + 4 * i);
            break;
        }
    if (tb.level == kNullLevel) {
        backtrack_to(s, 0);
        return GPSAT_UNSAT;
    }
    tb.branched = 1;
    tb.lit ^= 1;
    if (switch_literal) {
        backtrack_to(s, tb.level - 1);
        new_decision(s, tb);
        s.level++;
        Dec g = tb;
        g.formula = 0;
        graph_set(s, g);
    } else if (tb.level == lvl) {
        backtrack_to(s, tb.level - 1);
    } else {
        backtrack_to(s, tb.level);
    }
    return GPSAT_UNDEF;
}

// ---- conflict analysis: GraphAnalyzer's depth-first walk on an explicit stack
struct Learn {
    int32_t n;
    Dec hi, second;
};
__host__ __device__ void add_lit(RS& s, Learn& L, Dec d)
{
    s.lbuf[L.n++] = d.lit;
    d.lit ^= 1;
    if (d.level > L.hi.level) {
        L.second = L.hi;
        L.hi = d;
    } else if (d.level > L.second.level) {
        L.second = d;
    }
}
__host__ __device__ void analyze(RS& s, Learn& L)
{
    const int32_t root = s.n;
    int32_t sp = 0;
    // visit the conflict vertex
    s.vx_flag[root] = 1;
    if (s.vx_nb[root] == 0) {
        add_lit(s, L, get_dec(s.vx + 4 * root));
        return;
    }
    s.astack[0] = root;
    s.astack[1] = 0;
    sp = 1;
    while (sp > 0) {
        const int32_t v = s.astack[2 * (sp - 1)];
        const int32_t j = s.astack[2 * (sp - 1) + 1];
        if (j >= s.vx_nb[v]) {
            sp--;
            continue;
        }
        s.astack[2 * (sp - 1) + 1] = j + 1;
        const int32_t u = s.edges[(int64_t)v * s.ecap + j];
        const Dec nd = get_dec(s.vx + 4 * u);
        if (nd.level < s.vx[4 * v + 1]) {
            if (!s.vx_flag[u]) {
                add_lit(s, L, nd);
                s.vx_flag[u] = 1;
            }
        } else if (!s.vx_flag[u]) {
            s.vx_flag[u] = 1;
            if (s.vx_nb[u] == 0) {
                add_lit(s, L, nd);
            } else {
                s.astack[2 * sp] = u;
                s.astack[2 * sp + 1] = 0;
                sp++;
            }
        }
    }
}

// LearntClausesManager::learn_clause for the clause in the current slot (lstore slot 20)
__host__ __device__ void learn_clause(RS& s, int32_t id, int32_t len)
{
    int32_t* cur = s.lstore + kSlots * s.n;
    int32_t p;
    if (s.nlrep < kSlots) {
        p = s.nlrep++;
        const int32_t w = s.n_wc++;
        s.lrep_id[p] = id;
        s.lrep_wc[p] = w;
    } else {
        p = s.lpos;
        if (++s.lpos >= s.nlrep) s.lpos = 0;
        purge(s, s.lrep_wc[p]);        // reads the old clause's literals before they are replaced
        s.lrep_id[p] = id;
    }
    const int32_t w = s.lrep_wc[p];
    int32_t* dst = s.lstore + p * s.n;
    for (int32_t i = 0; i < len; i++) dst[i] = cur[i];
    s.wc_clause[w] = id;
    s.wc_start[w] = -1 - p * s.n;
    s.wc_len[w] = len;
    new_clause_checked(s, w);
}

// propagate_all_clauses: the walk, then every implication it found, in order
__host__ __device__ int32_t propagate_all(RS& s, const Dec& d)
{
    s.npend = 0;
    const int32_t confl = wl_new_decision(s, d.lit, d.formula);
    for (int32_t i = 0; i < s.npend; i++) {
        const int32_t* p = s.pend + 4 * i;
        s.pmark[p[0]] = 0;
        add_implication(s, Dec{p[0], p[1], 0, p[2]}, wc_lits(s, p[3]), s.wc_len[p[3]]);
    }
    s.npend = 0;
    return confl;
}

// The whole search of one instance (host and device code alike, so the same steps can be
// run on the CPU).
__host__ __device__ void solve_ref_instance(const RefArgs& a, int32_t inst)
{
    const InstMeta mi = a.meta[inst];
    const RefMeta rm = a.rmeta[inst];
    RS s;
    const int32_t n = mi.n, m = mi.m, nv = n > 0 ? n : 1;
    s.n = n;
    s.m = m;
    s.nw = rm.nw;
    s.ecap = rm.ecap;
    s.max_impl = rm.ecap - 1;
    s.olits = a.lits + mi.lit_base;
    s.ocoff = a.coff + mi.coff_base;
    int32_t* p = a.ws + rm.ws;
    const int32_t NW = rm.nw;
    s.wc_clause = p; p += NW;
    s.wc_w1 = p; p += NW;
    s.wc_w2 = p; p += NW;
    s.wc_start = p; p += NW;
    s.wc_len = p; p += NW;
    s.lstore = p; p += (kSlots + 1) * nv;
    s.lrep_id = p; p += kSlots;
    s.lrep_wc = p; p += kSlots;
    s.nd_val = p; p += 3 * NW;
    s.nd_next = p; p += 3 * NW;
    s.vlist = p; p += nv;
    s.st = p; p += nv;
    s.freev = p; p += nv + 1;
    s.decs = p; p += 4 * (nv + 1);
    s.imps = p; p += 4 * (nv + 1);
    s.vx = p; p += 4 * (nv + 1);
    s.vx_nb = p; p += nv + 1;
    s.vx_set = p; p += nv + 1;
    s.vx_flag = p; p += nv + 1;
    s.vidx = p; p += nv + 1;
    s.pend = p; p += 4 * (nv + 1);
    s.pmark = p; p += 2 * nv;
    s.frames = p; p += 6 * (nv + 2);
    s.snap = p; p += NW;
    s.lbuf = p; p += nv + 2;
    s.astack = p; p += 2 * (nv + 2);
    s.edges = p;
    // ---- initial state
    s.n_wc = 0;
    s.nlrep = 0;
    s.lpos = 0;
    s.nd_free = -1;
    for (int32_t k = 3 * NW - 1; k >= 0; k--) {
        s.nd_next[k] = s.nd_free;
        s.nd_free = k;
    }
    s.blocked = -1;
    s.nfree = 0;
    s.ndec = s.nimp = 0;
    s.level = 0;
    s.nvidx = 0;
    s.npend = 0;
    s.overflow = false;
    for (int32_t v = 0; v < nv; v++) {
        s.vlist[v] = -1;
        s.st[v] = -1;
    }
    for (int32_t l = 0; l < 2 * nv; l++) s.pmark[l] = 0;
    const int64_t vb = a.var_base[inst];
    for (int32_t v = 0; v <= n; v++) {
        put_dec(s.vx + 4 * v, Dec{2 * v + 1, -30, 0, 0});
        s.vx_nb[v] = 0;
        s.vx_set[v] = 0;
        s.vx_flag[v] = 0;
    }
    s.vx[4 * n + 1] = kIntMax;             // CONFLICT_VERTEX_DECISION_LEVEL
    for (int32_t v = 0; v < n; v++)
        if (!(a.dead && a.dead[vb + v])) free_add(s, v);
    // add_all_clauses: clauses of two or more literals, in order
    for (int32_t c = 0; c < m; c++) {
        const int32_t b = s.ocoff[c], len = s.ocoff[c + 1] - b;
        if (len < 2) continue;
        const int32_t w = s.n_wc++;
        s.wc_clause[w] = c;
        s.wc_start[w] = b;
        s.wc_len[w] = len;
        watch_clause(s, w, false);
    }
    s.next_id = m;
    int64_t st_dec = 0, st_conf = 0, st_learn = 0, st_rest = 0;
    int32_t iterations = 0, restart_thr = 100, restart_cnt = 0;
    int32_t status = GPSAT_UNDEF;
    if (s.nfree == 0) status = GPSAT_SAT;
    int32_t* cur = s.lstore + kSlots * nv;
    while (status == GPSAT_UNDEF) {
        Dec d{2 * s.freev[s.nfree - 1] + 1, s.level + 1, 0, 0};
        new_decision(s, d);
        s.level++;
        st_dec++;
        int32_t n_last = 0, pst = GPSAT_UNDEF;
        graph_set(s, d);
        int32_t confl = propagate_all(s, d);
        while (confl >= 0) {
            n_last++;
            st_conf++;
            if (s.level <= 0) {
                handle_backtrack(s, -1, true);
                pst = GPSAT_UNSAT;
                break;
            }
            const int32_t* cc = wc_lits(s, confl);
            for (int32_t i = 0; i < s.wc_len[confl]; i++) {
                if (!s.vx_set[n]) {
                    s.vx_set[n] = 1;
                    vidx_add(s, n);
                }
                s.vx[4 * n + 1] = s.level;
                s.vx[4 * n + 3] = 0;
                graph_edge(s, cc[i] >> 1, n);
            }
            Learn L;
            L.n = 0;
            L.hi = Dec{0, -1, 0, 0};
            L.second = Dec{0, -2, 0, 0};
            analyze(s, L);
            for (int32_t i = 0; i < s.nvidx; i++) s.vx_flag[s.vidx[i]] = 0;
            // the learned clause, in the current slot
            for (int32_t i = 0; i < L.n; i++) cur[i] = s.lbuf[i] ^ 1;
            const int32_t id = s.next_id++;
            st_learn++;
            if (L.n == 1) L.second.level = 0;
            Dec next{0, 0, 0, 0};
            int32_t cst;
            if (L.hi.level == 0) {
                handle_backtrack(s, 0, false);
                cst = GPSAT_UNSAT;
            } else if (L.n == 1) {
                cst = handle_backtrack(s, 1, false);
                next = Dec{cur[0], 0, 0, 1};
                graph_set(s, next);
                new_implication(s, next);
            } else if (L.second.level == L.hi.level) {
                cst = handle_backtrack(s, L.hi.level, true);
                if (s.ndec > 0) next = get_dec(s.decs + 4 * (s.ndec - 1));
                next.formula = 0;
            } else {
                cst = handle_backtrack(s, L.second.level + 1, false);
                next = L.hi;
                next.level = s.level;
                next.formula = 0;
                add_implication(s, next, cur, L.n);
            }
            if (L.n < s.max_impl && L.n > 1) learn_clause(s, id, L.n);
            if (cst != GPSAT_UNDEF) {
                pst = cst;
                break;
            }
            confl = propagate_all(s, next);
        }
        if (pst != GPSAT_UNDEF) {
            status = pst;
            break;
        }
        if (s.nfree == 0) {
            status = GPSAT_SAT;
            break;
        }
        restart_cnt += n_last;
        if (restart_cnt >= restart_thr) {
            restart_cnt = 0;
            restart_thr = (int32_t)((float)restart_thr * 1.3f);
            handle_backtrack(s, 0, false);
            st_rest++;
        }
        iterations++;
        if (a.max_iterations > 0 && iterations > a.max_iterations) break;
        if (s.overflow) break;
    }
    if (s.overflow && status == GPSAT_UNDEF) status = 4;
    a.status[inst] = status;
    int64_t* so = a.stats + (int64_t)inst * 5;
    so[0] = st_dec;
    so[1] = st_conf;
    so[2] = st_learn;
    so[3] = iterations;
    so[4] = st_rest;
    uint32_t* row = a.model_rows + mi.word_base;
    const int32_t W = words_per_row(n);
    for (int32_t wd = 0; wd < W; wd++) {
        uint32_t x = 0;
        if (status == GPSAT_SAT)
            for (int32_t b = 0; b < 32; b++) {
                const int32_t l = 32 * wd + b;
                if (l < 2 * n && lit_status(s, l) == GPSAT_SAT) x |= 1u << b;
            }
        row[wd] = x;
    }
}

__global__ __launch_bounds__(64) void k_solve_ref(RefArgs a)
{
    const int32_t inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst < a.n_inst) solve_ref_instance(a, inst);
}

// One instance per wavefront, run by its first lane: the searches of different instances
// diverge at almost every branch, so a wavefront of 64 instances runs them nearly one after
// another; one per wavefront puts as many independent searches in flight as the GPU holds
// wavefronts.
__global__ __launch_bounds__(256) void k_solve_ref_wave(RefArgs a)
{
    const int32_t inst = (int32_t)(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
    if (inst < a.n_inst && (threadIdx.x & 63) == 0) solve_ref_instance(a, inst);
}

}  // namespace

// Per-instance workspace ints of the trajectory mode (the slices carved in k_solve_ref).
int64_t solve_ref_ws_ints(int32_t n, int32_t m, int32_t ecap)
{
    const int64_t nv = n > 0 ? n : 1, NW = (int64_t)m + kSlots + 1;
    return 5 * NW + (kSlots + 1) * nv + 2 * kSlots + 6 * NW + 2 * nv + (nv + 1) + 12 * (nv + 1) + 4 * (nv + 1) +
           4 * (nv + 1) + 2 * nv + 6 * (nv + 2) + NW + (nv + 2) + 2 * (nv + 2) + (nv + 1) * (int64_t)ecap;
}

hipError_t launch_solve_ref(const RefArgs& a, hipStream_t st)
{
    if (a.n_inst == 0) return hipSuccess;
    if (tuning().solve_ref_per_wave)
        hipLaunchKernelGGL(k_solve_ref_wave, dim3((a.n_inst + 3) / 4), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(k_solve_ref, dim3((a.n_inst + 63) / 64), dim3(64), 0, st, a);
    return hipGetLastError();
}

}  // namespace gpsat

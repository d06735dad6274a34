/*
 * refsolve.c -- the reference's sequential solver, restated step for step for trajectory
 * parity (TEST INFRASTRUCTURE ONLY: tests/, smoke and bench's CPU baseline load it as the
 * checker; the product never does).
 *
 * Followed (file:line in the reference's src/):
 *   SATSolver::solve loop             SATSolver/SATSolver.cu:67-218
 *   DecisionMaker::decide             SATSolver/DecisionMaker.cu:45-86 (last free var, positive)
 *   VariablesStateHandler             SATSolver/VariablesStateHandler.cu:4-33 (free list),
 *                                     :113-206 (literal/clause status), :213-263 (reset order),
 *                                     :325-399 (new implication/decision, backtrack_to),
 *                                     :414-470 (is_unresolved, last_non_branched_decision),
 *                                     :560-588 (should_implicate_from_formula)
 *   Backtracker::handle_backtrack     SATSolver/Backtracker.cu:17-75
 *   ConflictAnalyzer                  ConflictAnalysis/ConflictAnalyzer.cu:34-171 (conflict,
 *                                     learning, asserting implication), :246-289 (propagate),
 *                                     :407-422 (add_implication), :494-523 (conflict vertex,
 *                                     learn_clause, restart)
 *   ConflictAnalyzerWithWatchedLits   ConflictAnalysis/ConflictAnalyzerWithWatchedLits.cu:113-139
 *   analyze_graph                     ConflictAnalysis/GraphAnalyzer.cu:31-138 (DFS from the
 *                                     conflict vertex; a neighbour below the *current vertex's*
 *                                     level joins the clause, others are recursed into)
 *   implication graph                 ConflictAnalysis/CUDAListGraph.cu:9-96,
 *                                     GraphStructure.cu:79-330 (per-var vertices, backward edges
 *                                     in insertion order, vertex removal drops the first edge to
 *                                     it from every other vertex, conflict vertex at index n)
 *   two watched literals              BCPStrategy/WatchedClausesList.cu:46-309 (new_decision:
 *                                     snapshot of the decided var's list, process_clause, then
 *                                     depth-first recursion into the implications found),
 *                                     :335-398 (new_clause, add_all_clauses)
 *   watch lists                       BCPStrategy/ClauseListStructure.cu:39-189 (per-variable
 *                                     lists, newest first; watch the two literals whose vars have
 *                                     the fewest watched clauses; blocked list; unblock_all
 *                                     re-watches unresolved blocked clauses on backtrack),
 *                                     :191-243 (purge_clause)
 *   learned clauses                   ClauseLearning/LearntClausesManager.cu:16-41,
 *                                     LearntClauseRepository.cu:15-33 (20 slots, round robin),
 *                                     WatchedClausesList.cu:33-46 (replace_clause)
 *   restarts                          Restarts/GeometricRestartsManager.cu:3-31 (100, x1.3)
 *
 * Where the reference reads out of bounds (a fresh learned clause blocked before it was ever
 * watched reads its watch index -1), the restatement does what that read leaves in effect:
 * the clause is in no var list, so only the blocked list changes.  Clauses of fewer than two
 * literals are never watched (add_all_clauses, WatchedClausesList.cu:386-395), so an empty
 * clause left by preprocessing is not seen by the search.
 */
#include "oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>

#define V_OF(l) ((l) >> 1)
#define NOT(l) ((l) ^ 1)
#define NULL_LEVEL (-234)      /* NULL_DECISION_LEVEL, Configs.cuh:24 */
#define LEARNT_SLOTS 20        /* MAX_LEARNT_CLAUSES_PER_THREAD, Configs.cuh:94 */

typedef struct { int32_t lit, level; uint8_t branched, formula; } dec_t;
typedef struct { dec_t d; int32_t clause; } impl_t;

typedef struct {
    int32_t n;
    /* clauses: id -> literals (originals 0..m-1, learned m, m+1, ...); identity = id */
    int32_t *clit; int64_t *cstart; int32_t *clen; int32_t ncl, capcl; int64_t nlit, caplit;
    /* watched-clause records (repository slots): clause id and the two watched indices */
    int32_t *wc_clause, *wc_w1, *wc_w2; int32_t n_wc;
    /* linked-list node pool shared by the per-var lists and the blocked list */
    int32_t *nd_val, *nd_next; int32_t nd_cap, nd_free;
    int32_t *vlist;          /* per var: head node (-1: empty) */
    int32_t blocked;         /* head node of the blocked list */
    /* variables: status 1 true / 0 false / -1 undef; ordered free list; decisions; implications */
    int8_t *st;
    int32_t *freev; int32_t nfree;
    dec_t *decs; int32_t ndec;
    dec_t *imps; int32_t nimp;
    int32_t level;
    /* implication graph: vertex v (var) or n (conflict); backward edges, insertion order */
    dec_t *vx; int32_t *vx_nb; uint8_t *vx_set, *vx_flag;
    int32_t *edges; int32_t ecap;             /* edges[v * ecap + k] = source vertex */
    int32_t *vidx; int32_t nvidx;             /* vertices_indices, insertion order */
    /* pending implications of one propagation and the literals among them */
    impl_t *pend; int32_t npend, cappend;
    uint8_t *pmark;                           /* pmark[lit]: lit is a pending implication */
    /* learned-clause repository */
    int32_t lrep[LEARNT_SLOTS]; int32_t nlrep, lpos;
    int32_t max_impl;
    int overflow;            /* a fixed capacity of the reference was exceeded */
} rs_t;

/* ---- node pool / linked lists (GPULinkedList: add_first, iterate, remove in place) */
static int32_t node_new(rs_t *s, int32_t val)
{
    int32_t k = s->nd_free;
    if (k < 0) {
        int32_t old = s->nd_cap;
        s->nd_cap = old * 2 + 64;
        s->nd_val = (int32_t *)realloc(s->nd_val, sizeof(int32_t) * s->nd_cap);
        s->nd_next = (int32_t *)realloc(s->nd_next, sizeof(int32_t) * s->nd_cap);
        for (int32_t i = s->nd_cap - 1; i >= old; i--) { s->nd_next[i] = s->nd_free; s->nd_free = i; }
        k = s->nd_free;
    }
    s->nd_free = s->nd_next[k];
    s->nd_val[k] = val;
    return k;
}
static void node_free(rs_t *s, int32_t k) { s->nd_next[k] = s->nd_free; s->nd_free = k; }
static void list_add_first(rs_t *s, int32_t *head, int32_t val)
{
    int32_t k = node_new(s, val);
    s->nd_next[k] = *head;
    *head = k;
}
static int32_t list_size(const rs_t *s, int32_t head)
{
    int32_t c = 0;
    for (int32_t k = head; k >= 0; k = s->nd_next[k]) c++;
    return c;
}
/* remove the first node whose watched clause is clause id c (ClauseListStructure::remove_clause) */
static int list_remove_clause(rs_t *s, int32_t *head, int32_t c)
{
    int32_t prev = -1;
    for (int32_t k = *head; k >= 0; prev = k, k = s->nd_next[k]) {
        if (s->wc_clause[s->nd_val[k]] == c) {
            if (prev < 0) *head = s->nd_next[k]; else s->nd_next[prev] = s->nd_next[k];
            node_free(s, k);
            return 1;
        }
    }
    return 0;
}

/* ---- variable state (VariablesStateHandler) */
static int lit_status(const rs_t *s, int32_t l)     /* 0 SAT, 1 UNSAT, 2 UNDEF */
{
    int8_t x = s->st[V_OF(l)];
    if (x < 0) return ORC_UNDEF;
    return ((l & 1) == x) ? ORC_SAT : ORC_UNSAT;
}
static int is_consistent(const rs_t *s, int32_t l) { return lit_status(s, l) != ORC_UNSAT; }
static void free_remove(rs_t *s, int32_t v)
{
    for (int32_t i = 0; i < s->nfree; i++)
        if (s->freev[i] == v) {
            memmove(s->freev + i, s->freev + i + 1, sizeof(int32_t) * (s->nfree - i - 1));
            s->nfree--;
            return;
        }
}
/* free_vars holds n_vars entries; GPUVec::add drops an add to a full list */
static void free_add(rs_t *s, int32_t v)
{
    if (s->nfree < s->n) s->freev[s->nfree++] = v;
}
static void new_implication(rs_t *s, dec_t d)
{
    s->imps[s->nimp++] = d;
    free_remove(s, V_OF(d.lit));
    s->st[V_OF(d.lit)] = (int8_t)(d.lit & 1);
}
static void new_decision(rs_t *s, dec_t d)
{
    s->decs[s->ndec++] = d;
    free_remove(s, V_OF(d.lit));
    s->st[V_OF(d.lit)] = (int8_t)(d.lit & 1);
}
static int is_formula_implied(const rs_t *s, int32_t v)
{
    for (int32_t i = 0; i < s->nimp; i++)
        if (V_OF(s->imps[i].lit) == v) return s->imps[i].formula;
    return 0;
}
static int should_implicate_from_formula(const rs_t *s, int32_t c, int32_t unit)
{
    const int32_t *cl = s->clit + s->cstart[c];
    int32_t count = 0;
    for (int32_t i = 0; i < s->clen[c]; i++)
        if (V_OF(cl[i]) == V_OF(unit) || is_formula_implied(s, V_OF(cl[i]))) count++;
    return count == s->clen[c];
}
static int is_unresolved(const rs_t *s, int32_t c)
{
    const int32_t *cl = s->clit + s->cstart[c];
    int32_t n = s->clen[c], n_unsat = 0;
    if (n == 1) return 0;
    for (int32_t i = 0; i < n; i++) {
        int ls = lit_status(s, cl[i]);
        if (ls == ORC_SAT) return 0;
        if (ls == ORC_UNSAT) n_unsat++;
    }
    return n_unsat < n - 1;
}

/* ---- watch structure (ClauseListStructure) */
static void watch_clause(rs_t *s, int32_t w, int test_consistency)
{
    const int32_t c = s->wc_clause[w];
    const int32_t *cl = s->clit + s->cstart[c];
    int32_t sv = INT_MAX - 1, ssv = INT_MAX, sm = 0, ssm = 0;
    for (int32_t i = 0; i < s->clen[c]; i++) {
        if (test_consistency && !is_consistent(s, cl[i])) continue;
        int32_t size = list_size(s, s->vlist[V_OF(cl[i])]);
        if (size < sv) { ssm = sm; ssv = sv; sv = size; sm = i; }
        else if (size < ssv) { ssv = size; ssm = i; }
    }
    s->wc_w1[w] = sm;
    s->wc_w2[w] = ssm;
    list_add_first(s, &s->vlist[V_OF(cl[sm])], w);
    list_add_first(s, &s->vlist[V_OF(cl[ssm])], w);
}
static void block_clause(rs_t *s, int32_t w)
{
    const int32_t c = s->wc_clause[w];
    const int32_t *cl = s->clit + s->cstart[c];
    if (s->wc_w1[w] >= 0) list_remove_clause(s, &s->vlist[V_OF(cl[s->wc_w1[w]])], c);
    if (s->wc_w2[w] >= 0) list_remove_clause(s, &s->vlist[V_OF(cl[s->wc_w2[w]])], c);
    list_add_first(s, &s->blocked, w);
}
static void switch_watch(rs_t *s, int32_t old_i, int32_t new_i, int32_t w)
{
    const int32_t c = s->wc_clause[w];
    const int32_t *cl = s->clit + s->cstart[c];
    if (s->wc_w1[w] == old_i) s->wc_w1[w] = new_i; else s->wc_w2[w] = new_i;
    list_remove_clause(s, &s->vlist[V_OF(cl[old_i])], c);
    list_add_first(s, &s->vlist[V_OF(cl[new_i])], w);
}
static void unblock_all(rs_t *s)
{
    int32_t prev = -1;
    for (int32_t k = s->blocked; k >= 0;) {
        int32_t nxt = s->nd_next[k], w = s->nd_val[k];
        if (is_unresolved(s, s->wc_clause[w])) {
            watch_clause(s, w, 1);
            if (prev < 0) s->blocked = nxt; else s->nd_next[prev] = nxt;
            node_free(s, k);
        } else {
            prev = k;
        }
        k = nxt;
    }
}
/* new_clause(..., check_status = true) into slot w (WatchedClausesList.cu:358-375) */
static void new_clause_checked(rs_t *s, int32_t w, int32_t c)
{
    const int32_t *cl = s->clit + s->cstart[c];
    int32_t n = s->clen[c], n_unsat = 0, unit = -1, sat = 0;
    s->wc_clause[w] = c;
    for (int32_t i = 0; i < n && !sat; i++) {
        int ls = lit_status(s, cl[i]);
        if (ls == ORC_SAT) sat = 1;
        else if (ls == ORC_UNSAT) n_unsat++;
        else unit = cl[i];
    }
    if (sat || n_unsat != n - 1) unit = -1;
    if (!sat && (n_unsat == n || unit >= 0)) {
        /* the clause is in no var list yet, so blocking it only adds it to the blocked list
         * (the reference looks it up in the lists of the vars its slot's previous watch
         * indices name and finds nothing) */
        s->wc_w1[w] = s->wc_w2[w] = -1;
        block_clause(s, w);
    } else {
        watch_clause(s, w, 0);
    }
}
/* purge_clause (ClauseListStructure.cu:191-243) */
static void purge_clause(rs_t *s, int32_t c)
{
    int32_t prev = -1;
    for (int32_t k = s->blocked; k >= 0; prev = k, k = s->nd_next[k])
        if (s->wc_clause[s->nd_val[k]] == c) {
            if (prev < 0) s->blocked = s->nd_next[k]; else s->nd_next[prev] = s->nd_next[k];
            node_free(s, k);
            return;
        }
    const int32_t *cl = s->clit + s->cstart[c];
    int32_t other = -1;
    for (int32_t i = 0; i < s->clen[c] && other < 0; i++) {
        int32_t v = V_OF(cl[i]);
        int32_t pv = -1;
        for (int32_t k = s->vlist[v]; k >= 0; pv = k, k = s->nd_next[k]) {
            int32_t w = s->nd_val[k];
            if (s->wc_clause[w] == c) {
                int32_t v1 = V_OF(cl[s->wc_w1[w]]), v2 = V_OF(cl[s->wc_w2[w]]);
                other = (v1 == v) ? v2 : v1;
                if (pv < 0) s->vlist[v] = s->nd_next[k]; else s->nd_next[pv] = s->nd_next[k];
                node_free(s, k);
                break;
            }
        }
    }
    if (other >= 0) list_remove_clause(s, &s->vlist[other], c);
}

/* ---- pending implications of one propagation (WatchedClausesList) */
static int currently_consistent(const rs_t *s, int32_t l) { return is_consistent(s, l) && !s->pmark[NOT(l)]; }
static int pending_status(const rs_t *s, int32_t l)
{
    int ls = lit_status(s, l);
    if (ls != ORC_UNDEF) return ls;
    if (s->pmark[l]) return ORC_SAT;
    if (s->pmark[NOT(l)]) return ORC_UNSAT;
    return ORC_UNDEF;
}
static void handle_implication(rs_t *s, int32_t w, int32_t unit, int from_formula)
{
    if (pending_status(s, unit) == ORC_UNDEF) {
        if (s->npend == s->cappend) {
            s->cappend = s->cappend * 2 + 16;
            s->pend = (impl_t *)realloc(s->pend, sizeof(impl_t) * s->cappend);
        }
        impl_t *p = &s->pend[s->npend++];
        p->d.lit = unit;
        p->d.level = s->level;
        p->d.branched = 0;
        p->d.formula = (uint8_t)(from_formula ? should_implicate_from_formula(s, s->wc_clause[w], unit) : 0);
        p->clause = s->wc_clause[w];
        s->pmark[unit] = 1;
    }
    block_clause(s, w);
}
/* process_clause (WatchedClausesList.cu:106-221); 1 on a conflict */
static int process_clause(rs_t *s, int32_t w, int from_formula)
{
    const int32_t c = s->wc_clause[w];
    const int32_t *cl = s->clit + s->cstart[c];
    const int32_t n = s->clen[c];
    const int32_t l1 = cl[s->wc_w1[w]], l2 = cl[s->wc_w2[w]];
    const int c1 = currently_consistent(s, l1), c2 = currently_consistent(s, l2);
    if (c1 && c2) return 0;
    int32_t pending = (!c1 && !c2) ? 2 : 1, o1 = -1, o2 = -1;
    for (int32_t i = 0; pending > 0 && i < n; i++) {
        int32_t l = cl[i];
        if (l != l1 && l != l2 && currently_consistent(s, l)) {
            if (o1 == -1) o1 = i; else o2 = i;
            pending--;
        }
    }
    if (c1 || c2) {
        if (o1 == -1) handle_implication(s, w, cl[c1 ? s->wc_w1[w] : s->wc_w2[w]], from_formula);
        else switch_watch(s, !c1 ? s->wc_w1[w] : s->wc_w2[w], o1, w);
        return 0;
    }
    if (o1 != -1 && o2 != -1) {
        switch_watch(s, s->wc_w1[w], o1, w);
        switch_watch(s, s->wc_w2[w], o2, w);
        return 0;
    }
    if (o1 != -1 || o2 != -1) {
        handle_implication(s, w, cl[o1 != -1 ? o1 : o2], from_formula);
        return 0;
    }
    block_clause(s, w);     /* handle_conflict */
    return 1;
}
/* new_decision (WatchedClausesList.cu:46-104): the decided var's list as it stands, then
 * depth-first into each implication this call found.  Returns the conflicting clause or -1. */
static int32_t wl_new_decision(rs_t *s, int32_t lit, int from_formula)
{
    int32_t cnt = list_size(s, s->vlist[V_OF(lit)]);
    int32_t *snap = (int32_t *)malloc(sizeof(int32_t) * (cnt + 1));
    int32_t i = 0;
    for (int32_t k = s->vlist[V_OF(lit)]; k >= 0; k = s->nd_next[k]) snap[i++] = s->nd_val[k];
    int32_t start = s->npend;
    for (i = 0; i < cnt; i++)
        if (process_clause(s, snap[i], from_formula)) {
            int32_t c = s->wc_clause[snap[i]];
            free(snap);
            return c;
        }
    free(snap);
    int32_t found = s->npend - start;
    for (i = 0; i < found; i++) {
        impl_t p = s->pend[start + i];
        int32_t c = wl_new_decision(s, p.d.lit, p.d.formula);
        if (c >= 0) return c;
    }
    return -1;
}

/* ---- implication graph (CUDAListGraph / GraphStructure) */
static void vidx_remove(rs_t *s, int32_t v)
{
    for (int32_t i = 0; i < s->nvidx; i++)
        if (s->vidx[i] == v) {
            memmove(s->vidx + i, s->vidx + i + 1, sizeof(int32_t) * (s->nvidx - i - 1));
            s->nvidx--;
            return;
        }
}
/* vertices_indices holds n + 1 entries (GraphStructure.cu:5); an add to a full list is
 * dropped, as GPUVec::add drops it */
static void vidx_add(rs_t *s, int32_t v)
{
    if (s->nvidx < s->n + 1) s->vidx[s->nvidx++] = v;
}
static void graph_set(rs_t *s, dec_t d)
{
    int32_t v = V_OF(d.lit);
    s->vx[v] = d;
    s->vx_nb[v] = 0;
    s->vx_set[v] = 1;
    vidx_add(s, v);
}
static void graph_edge(rs_t *s, int32_t src, int32_t dst)
{
    if (s->vx_nb[dst] >= s->ecap) { s->overflow = 1; return; }
    s->edges[(int64_t)dst * s->ecap + s->vx_nb[dst]++] = src;
}
static void graph_remove_vertex(rs_t *s, int32_t v)
{
    s->vx_set[v] = 0;
    s->vx_flag[v] = 0;
    s->vx_nb[v] = 0;
    vidx_remove(s, v);
    for (int32_t i = 0; i < s->nvidx; i++) {
        int32_t o = s->vidx[i];
        int32_t *e = s->edges + (int64_t)o * s->ecap;
        for (int32_t j = 0; j < s->vx_nb[o]; j++)
            if (e[j] == v) {
                memmove(e + j, e + j + 1, sizeof(int32_t) * (s->vx_nb[o] - j - 1));
                s->vx_nb[o]--;
                break;
            }
    }
}
static void graph_backtrack(rs_t *s, int32_t lvl)
{
    for (int32_t i = 0; i < s->nvidx; i++) {
        int32_t v = s->vidx[i];
        if (s->vx[v].level > lvl && !s->vx[v].formula) {
            graph_remove_vertex(s, v);
            i--;
        }
    }
}
/* ConflictAnalyzer::add_implication (ConflictAnalyzer.cu:407-422) */
static void add_implication(rs_t *s, dec_t d, int32_t c)
{
    graph_set(s, d);
    const int32_t *cl = s->clit + s->cstart[c];
    for (int32_t i = 0; i < s->clen[c]; i++)
        if (cl[i] != d.lit) graph_edge(s, V_OF(cl[i]), V_OF(d.lit));
    new_implication(s, d);
}

/* ---- backtracking (Backtracker, VariablesStateHandler::backtrack_to) */
static void backtrack_to(rs_t *s, int32_t lvl)
{
    s->level = lvl;
    for (int32_t i = 0; i < s->nimp; i++)
        if (s->imps[i].level > lvl) {
            int32_t v = V_OF(s->imps[i].lit);
            memmove(s->imps + i, s->imps + i + 1, sizeof(dec_t) * (s->nimp - i - 1));
            s->nimp--;
            free_add(s, v);
            s->st[v] = -1;
            i--;
        }
    for (int32_t i = 0; i < s->ndec; i++)
        if (s->decs[i].level > lvl) {
            int32_t v = V_OF(s->decs[i].lit);
            memmove(s->decs + i, s->decs + i + 1, sizeof(dec_t) * (s->ndec - i - 1));
            s->ndec--;
            free_add(s, v);
            s->st[v] = -1;
            i--;
        }
    graph_backtrack(s, lvl);
    unblock_all(s);
}
/* Backtracker::handle_backtrack (Backtracker.cu:37-75); returns ORC_UNSAT or ORC_UNDEF */
static int handle_backtrack(rs_t *s, int32_t lvl, int switch_literal)
{
    dec_t tb;
    tb.level = NULL_LEVEL;
    for (int32_t i = s->ndec - 1; i >= 0; i--)
        if (!s->decs[i].branched && s->decs[i].level <= lvl) { tb = s->decs[i]; break; }
    if (tb.level == NULL_LEVEL) {
        backtrack_to(s, 0);
        return ORC_UNSAT;
    }
    tb.branched = 1;
    tb.lit = NOT(tb.lit);
    if (switch_literal) {
        backtrack_to(s, tb.level - 1);
        new_decision(s, tb);
        s->level++;
        dec_t g = tb;
        g.formula = 0;
        graph_set(s, g);
    } else if (tb.level == lvl) {
        backtrack_to(s, tb.level - 1);
    } else {
        backtrack_to(s, tb.level);
    }
    return ORC_UNDEF;
}

/* ---- conflict analysis (GraphAnalyzer.cu) */
typedef struct { int32_t *lits; int32_t n; dec_t hi, second; } learn_t;
static void add_lit(learn_t *L, dec_t d)
{
    L->lits[L->n++] = d.lit;
    d.lit = NOT(d.lit);
    if (d.level > L->hi.level) { L->second = L->hi; L->hi = d; }
    else if (d.level > L->second.level) L->second = d;
}
static void conflicting_assignment(rs_t *s, int32_t v, learn_t *L)
{
    if (s->vx_flag[v]) return;
    s->vx_flag[v] = 1;
    const dec_t cur = s->vx[v];
    if (s->vx_nb[v] == 0) { add_lit(L, cur); return; }
    for (int32_t j = 0; j < s->vx_nb[v]; j++) {
        int32_t u = s->edges[(int64_t)v * s->ecap + j];
        dec_t nd = s->vx[u];
        if (nd.level < cur.level) {
            if (!s->vx_flag[u]) { add_lit(L, nd); s->vx_flag[u] = 1; }
        } else {
            conflicting_assignment(s, u, L);
        }
    }
}

/* store a learned clause's literals; returns its clause id */
static int32_t clause_add(rs_t *s, const int32_t *l, int32_t n)
{
    if (s->nlit + n > s->caplit) {
        while (s->nlit + n > s->caplit) s->caplit = s->caplit * 2 + 64;
        s->clit = (int32_t *)realloc(s->clit, sizeof(int32_t) * s->caplit);
    }
    if (s->ncl == s->capcl) {
        s->capcl = s->capcl * 2 + 16;
        s->cstart = (int64_t *)realloc(s->cstart, sizeof(int64_t) * s->capcl);
        s->clen = (int32_t *)realloc(s->clen, sizeof(int32_t) * s->capcl);
    }
    memcpy(s->clit + s->nlit, l, sizeof(int32_t) * n);
    s->cstart[s->ncl] = s->nlit;
    s->clen[s->ncl] = n;
    s->nlit += n;
    return s->ncl++;
}

/* LearntClausesManager::learn_clause + LearntClauseRepository::add_clause */
static void learn_clause(rs_t *s, int32_t c)
{
    if (s->nlrep < LEARNT_SLOTS) {
        s->lrep[s->nlrep++] = c;
        int32_t w = s->n_wc++;
        new_clause_checked(s, w, c);
    } else {
        int32_t old = s->lrep[s->lpos];
        s->lrep[s->lpos] = c;
        if (++s->lpos >= s->nlrep) s->lpos = 0;
        purge_clause(s, old);
        for (int32_t w = s->n_wc - 1; w >= 0; w--)
            if (s->wc_clause[w] == old) { new_clause_checked(s, w, c); break; }
    }
}

/* ConflictAnalyzerWithWatchedLits::propagate_all_clauses: returns the conflicting clause or -1 */
static int32_t propagate_all(rs_t *s, dec_t d)
{
    s->npend = 0;
    int32_t confl = wl_new_decision(s, d.lit, d.formula);
    for (int32_t i = 0; i < s->npend; i++) {
        s->pmark[s->pend[i].d.lit] = 0;
        add_implication(s, s->pend[i].d, s->pend[i].clause);
    }
    s->npend = 0;
    return confl;
}

int orc_solve_ref(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                  const uint8_t *dead, int32_t max_iterations, int8_t *model, int64_t *stats)
{
    rs_t S, *s = &S;
    memset(s, 0, sizeof(*s));
    const int32_t nv = n > 0 ? n : 1;
    s->n = n;
    s->caplit = (clause_off[m] - clause_off[0]) * 2 + 1024;
    s->clit = (int32_t *)malloc(sizeof(int32_t) * s->caplit);
    s->capcl = m + 64;
    s->cstart = (int64_t *)malloc(sizeof(int64_t) * s->capcl);
    s->clen = (int32_t *)malloc(sizeof(int32_t) * s->capcl);
    int32_t largest = 0;
    for (int32_t c = 0; c < m; c++) {
        int32_t len = (int32_t)(clause_off[c + 1] - clause_off[c]);
        if (len > largest) largest = len;
        clause_add(s, lits + clause_off[c], len);
    }
    /* max_implication_per_var = max(largest clause, MIN_IMPLICATION_PER_VAR) (main.cu:148) */
    s->max_impl = largest > 100 ? largest : 100;
    s->ecap = s->max_impl + 1;
    const int32_t wcap = m + 1 + LEARNT_SLOTS;
    s->wc_clause = (int32_t *)malloc(sizeof(int32_t) * wcap);
    s->wc_w1 = (int32_t *)malloc(sizeof(int32_t) * wcap);
    s->wc_w2 = (int32_t *)malloc(sizeof(int32_t) * wcap);
    s->nd_free = -1;
    s->vlist = (int32_t *)malloc(sizeof(int32_t) * nv);
    for (int32_t v = 0; v < nv; v++) s->vlist[v] = -1;
    s->blocked = -1;
    s->st = (int8_t *)malloc(nv);
    memset(s->st, -1, nv);
    s->freev = (int32_t *)malloc(sizeof(int32_t) * (nv + 1));
    s->decs = (dec_t *)malloc(sizeof(dec_t) * (nv + 1));
    s->imps = (dec_t *)malloc(sizeof(dec_t) * (nv + 1));
    s->vx = (dec_t *)calloc(nv + 1, sizeof(dec_t));
    s->vx_nb = (int32_t *)calloc(nv + 1, sizeof(int32_t));
    s->vx_set = (uint8_t *)calloc(nv + 1, 1);
    s->vx_flag = (uint8_t *)calloc(nv + 1, 1);
    s->edges = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv + 1) * s->ecap);
    s->vidx = (int32_t *)malloc(sizeof(int32_t) * (2 * nv + 8));
    s->pmark = (uint8_t *)calloc(2 * nv, 1);
    for (int32_t v = 0; v < n; v++) {
        s->vx[v].lit = 2 * v + 1;       /* vertex initialisation (GraphStructure.cu:33-42) */
        s->vx[v].level = -30;
        if (!(dead && dead[v])) free_add(s, v);
    }
    s->vx[n].level = INT_MAX;           /* CONFLICT_VERTEX_DECISION_LEVEL */
    /* add_all_clauses: every clause of two or more literals, in order */
    for (int32_t c = 0; c < m; c++) {
        if (s->clen[c] < 2) continue;
        int32_t w = s->n_wc++;
        s->wc_clause[w] = c;
        watch_clause(s, w, 0);
    }
    int64_t st_dec = 0, st_conf = 0, st_learn = 0, st_rest = 0;
    int32_t iterations = 0, restart_thr = 100, restart_cnt = 0;
    int status = ORC_UNDEF;
    int32_t *lbuf = (int32_t *)malloc(sizeof(int32_t) * (nv + 2));
    if (s->nfree == 0) status = ORC_SAT;
    while (status == ORC_UNDEF) {
        /* DecisionMaker::decide: last free var, positive, level + 1 */
        dec_t d;
        d.lit = 2 * s->freev[s->nfree - 1] + 1;
        d.level = s->level + 1;
        d.branched = 0;
        d.formula = 0;
        new_decision(s, d);
        s->level++;
        st_dec++;
        /* ConflictAnalyzer::propagate */
        int32_t n_last = 0;
        graph_set(s, d);
        int32_t confl = propagate_all(s, d);
        int pst = ORC_UNDEF;
        while (confl >= 0) {
            n_last++;
            st_conf++;
            /* handle_conflict_with_clause_learning */
            if (s->level <= 0) {
                handle_backtrack(s, -1, 1);
                pst = ORC_UNSAT;
                break;
            }
            const int32_t *cc = s->clit + s->cstart[confl];
            for (int32_t i = 0; i < s->clen[confl]; i++) {
                /* link_with_conflict (GraphStructure.cu:308-330) */
                if (!s->vx_set[n]) { s->vx_set[n] = 1; vidx_add(s, n); }
                s->vx[n].level = s->level;
                s->vx[n].formula = 0;
                graph_edge(s, V_OF(cc[i]), n);
            }
            learn_t L;
            L.lits = lbuf;
            L.n = 0;
            L.hi.level = -1;
            L.hi.lit = 0;
            L.second.level = -2;
            L.second.lit = 0;
            conflicting_assignment(s, n, &L);
            if (L.n > nv + 1) s->overflow = 1;
            for (int32_t i = 0; i < s->nvidx; i++) s->vx_flag[s->vidx[i]] = 0;
            for (int32_t i = 0; i < L.n; i++) lbuf[i] = NOT(lbuf[i]);
            const int32_t lc = clause_add(s, lbuf, L.n);
            st_learn++;
            if (L.n == 1) L.second.level = 0;
            /* implicate_and_backtrack_from_learnt_clause (ConflictAnalyzer.cu:105-151) */
            dec_t next;
            int cst;
            if (L.hi.level == 0) {
                handle_backtrack(s, 0, 0);
                cst = ORC_UNSAT;
            } else if (L.n == 1) {
                cst = handle_backtrack(s, 1, 0);
                next.lit = s->clit[s->cstart[lc]];
                next.level = 0;
                next.branched = 0;
                next.formula = 1;
                graph_set(s, next);
                new_implication(s, next);
            } else if (L.second.level == L.hi.level) {
                cst = handle_backtrack(s, L.hi.level, 1);
                next = s->ndec > 0 ? s->decs[s->ndec - 1] : next;
                next.formula = 0;
            } else {
                cst = handle_backtrack(s, L.second.level + 1, 0);
                next = L.hi;
                next.level = s->level;
                next.formula = 0;
                add_implication(s, next, lc);
            }
            /* handle_learnt_clause: kept when 2 <= size < max_implication_per_var */
            if (L.n < s->max_impl && L.n > 1) learn_clause(s, lc);
            if (cst != ORC_UNDEF) {
                pst = cst;
                break;
            }
            confl = propagate_all(s, next);
        }
        if (pst != ORC_UNDEF) { status = pst; break; }
        if (s->nfree == 0) { status = ORC_SAT; break; }
        /* GeometricRestartsManager, SATSolver.cu:170-178; ConflictAnalyzer::restart */
        restart_cnt += n_last;
        if (restart_cnt >= restart_thr) {
            restart_cnt = 0;
            restart_thr = (int32_t)((float)restart_thr * 1.3f);
            handle_backtrack(s, 0, 0);
            st_rest++;
        }
        iterations++;
        if (max_iterations > 0 && iterations > max_iterations) break;
        if (s->overflow) break;
    }
    if (s->overflow && status == ORC_UNDEF) status = 4;
    if (model) for (int32_t v = 0; v < n; v++) model[v] = status == ORC_SAT ? s->st[v] : -1;
    if (stats) { stats[0] = st_dec; stats[1] = st_conf; stats[2] = st_learn; stats[3] = iterations; stats[4] = st_rest; }
    free(s->clit); free(s->cstart); free(s->clen); free(s->wc_clause); free(s->wc_w1); free(s->wc_w2);
    free(s->nd_val); free(s->nd_next); free(s->vlist); free(s->st); free(s->freev); free(s->decs); free(s->imps);
    free(s->vx); free(s->vx_nb); free(s->vx_set); free(s->vx_flag); free(s->edges); free(s->vidx); free(s->pend);
    free(s->pmark); free(lbuf);
    return status;
}

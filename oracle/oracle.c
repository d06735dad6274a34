/*
 * oracle.c -- CPU restatement of the gpupsat solve path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
 * library, and only as the checker (or, for the baseline, as the timed CPU
 * restatement of the reference's sequential solver).  The product never links it.
 * Every function cites the reference file:line whose behaviour it restates.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define LIT_VAR(l) ((l) >> 1)
#define LIT_NEG(l) ((l) ^ 1)

static inline int row_bit(const uint32_t *row, int32_t l) { return (row[l >> 5] >> (l & 31)) & 1u; }
static inline void row_set(uint32_t *row, int32_t l) { row[l >> 5] |= 1u << (l & 31); }
static inline int32_t words_per_row(int32_t n) { return (2 * n + 31) / 32; }

/* ------------------------------------------------------------------------------------------
 * a4: VariablesStateHandler::clause_status (VariablesStateHandler.cu:180-206).
 * SAT as soon as one literal is true; otherwise count false literal occurrences;
 * the unit literal is the last non-false literal and is reported iff exactly
 * n-1 literals are false; UNSAT iff all n are false. */
int orc_clause_status(const int32_t *lits, int32_t n, const uint32_t *row, int32_t *unit_lit)
{
    int32_t nfalse = 0, last_undef = -1;
    for (int32_t i = 0; i < n; i++) {
        int32_t l = lits[i];
        if (row_bit(row, l)) { *unit_lit = -1; return ORC_SAT; }
        if (row_bit(row, LIT_NEG(l))) nfalse++;
        else last_undef = l;
    }
    *unit_lit = (nfalse == n - 1) ? last_undef : -1;
    return nfalse == n ? ORC_UNSAT : ORC_UNDEF;
}

void orc_eval_batch(int32_t n_inst, const int32_t *lits, const int64_t *clause_off,
                    const int64_t *inst_clause_off, const int32_t *n_vars,
                    const uint32_t *rows, int32_t A, int32_t *counts)
{
    int64_t wbase = 0;
    for (int32_t i = 0; i < n_inst; i++) {
        int32_t W = words_per_row(n_vars[i]);
        for (int32_t a = 0; a < A; a++) {
            const uint32_t *row = rows + (wbase * A) + (int64_t)a * W;
            int32_t c_sat = 0, c_unsat = 0, c_unit = 0;
            for (int64_t c = inst_clause_off[i]; c < inst_clause_off[i + 1]; c++) {
                int32_t u;
                int st = orc_clause_status(lits + clause_off[c], (int32_t)(clause_off[c + 1] - clause_off[c]), row, &u);
                if (st == ORC_SAT) c_sat++;
                else if (st == ORC_UNSAT) c_unsat++;
                else if (u >= 0) c_unit++;
            }
            int32_t *o = counts + ((int64_t)i * A + a) * 3;
            o[0] = c_sat; o[1] = c_unsat; o[2] = c_unit;
        }
        wbase += W;
    }
}

/* ------------------------------------------------------------------------------------------
 * a5: BCP fixpoint.  The reference's watched-literal BCP (WatchedClausesList.cu:46-309,
 * driven per trail literal as in WatchedClausesListTester.cu:22-109) reaches the
 * textbook unit-propagation closure of the trail; the closure does not depend on the
 * order in which trail literals are applied, and a conflict arises in some order iff it
 * arises in every order.  This is the plain closure by repeated clause passes. */
int orc_up_closure(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                   const uint32_t *in_row, uint32_t *out_row, int32_t *props)
{
    int32_t W = words_per_row(n);
    uint32_t *cur = (uint32_t *)malloc(sizeof(uint32_t) * (W ? W : 1));
    memcpy(cur, in_row, sizeof(uint32_t) * W);
    int conflict = 0;
    for (int32_t v = 0; v < n && !conflict; v++)
        if (row_bit(cur, 2 * v) && row_bit(cur, 2 * v + 1)) conflict = 1;
    int changed = 1;
    while (changed && !conflict) {
        changed = 0;
        for (int32_t c = 0; c < m && !conflict; c++) {
            int32_t u;
            int st = orc_clause_status(lits + clause_off[c], (int32_t)(clause_off[c + 1] - clause_off[c]), cur, &u);
            if (st == ORC_UNSAT) conflict = 1;
            else if (st == ORC_UNDEF && u >= 0) {
                if (row_bit(cur, LIT_NEG(u))) conflict = 1; /* cannot happen: u is non-false */
                row_set(cur, u);
                changed = 1;
            }
        }
    }
    if (conflict) {
        memcpy(out_row, in_row, sizeof(uint32_t) * W);
        *props = 0;
    } else {
        int32_t p = 0;
        for (int32_t w = 0; w < W; w++) p += __builtin_popcount(cur[w] & ~in_row[w]);
        memcpy(out_row, cur, sizeof(uint32_t) * W);
        *props = p;
    }
    free(cur);
    return conflict;
}

void orc_bcp_batch(int32_t n_inst, const int32_t *lits, const int64_t *clause_off,
                   const int64_t *inst_clause_off, const int32_t *n_vars,
                   const uint32_t *rows, int32_t A, uint32_t *out_rows,
                   int32_t *conflict, int32_t *props)
{
    int64_t wbase = 0;
    for (int32_t i = 0; i < n_inst; i++) {
        int32_t W = words_per_row(n_vars[i]);
        int64_t c0 = inst_clause_off[i];
        int32_t m = (int32_t)(inst_clause_off[i + 1] - c0);
        for (int32_t a = 0; a < A; a++) {
            int64_t off = wbase * A + (int64_t)a * W;
            conflict[(int64_t)i * A + a] = orc_up_closure(lits, clause_off + c0, m, n_vars[i], rows + off,
                                                          out_rows + off, props + (int64_t)i * A + a);
        }
        wbase += W;
    }
}

/* ------------------------------------------------------------------------------------------
 * a3: preprocessing.  RepeatedLiteralsRemover::process (RepeatedLiteralsRemover.cu:25-62):
 * keep the first occurrence of each literal, drop a clause holding both polarities of a
 * var; SAT if no clause is left.  UnaryClausesRemover::process (UnaryClausesRemover.cu:13-173):
 * unit clauses become solved literals (a duplicate unit stays in the formula, a
 * complementary one sets UNSAT); then passes over the formula in order: a clause with a
 * solved literal is erased, a clause left with one non-false literal adds that literal
 * and is erased, an all-false clause sets UNSAT, other clauses lose their false
 * literals; passes repeat while the solved set grows; SAT if no clause is left. */
typedef struct { int32_t *l; int32_t n; } pclause;

int orc_preprocess(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                   int32_t *out_lits, int64_t *out_clause_off, int32_t *out_m,
                   int32_t *solved, int32_t *n_solved)
{
    pclause *f = (pclause *)malloc(sizeof(pclause) * (m ? m : 1));
    int8_t *stat = (int8_t *)malloc(n ? n : 1);
    int32_t fm = 0;
    int status = ORC_UNDEF;
    *n_solved = 0;
    /* RepeatedLiteralsRemover */
    for (int32_t c = 0; c < m; c++) {
        int32_t len = (int32_t)(clause_off[c + 1] - clause_off[c]);
        int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (len ? len : 1));
        int32_t k = 0, taut = 0;
        memset(stat, 0, n);
        for (int32_t i = 0; i < len && !taut; i++) {
            int32_t l = lits[clause_off[c] + i];
            int8_t s = (l & 1) ? 1 : 2;
            if (stat[LIT_VAR(l)] == 0) { stat[LIT_VAR(l)] = s; buf[k++] = l; }
            else if (stat[LIT_VAR(l)] != s) taut = 1;
        }
        if (taut) { free(buf); continue; }
        f[fm].l = buf; f[fm].n = k; fm++;
    }
    free(stat);
    if (fm == 0) status = ORC_SAT;
    if (status == ORC_UNDEF) {
        /* process_unary_clauses: remove_if over the formula, order preserved */
        int32_t w = 0;
        for (int32_t c = 0; c < fm; c++) {
            int removed = 0;
            if (f[c].n == 1) {
                int32_t l = f[c].l[0];
                int dup = 0;
                for (int32_t i = 0; i < *n_solved; i++) {
                    if (solved[i] == l) { dup = 1; break; }
                    if (solved[i] == LIT_NEG(l)) { status = ORC_UNSAT; dup = 1; break; }
                }
                if (!dup) { solved[(*n_solved)++] = l; removed = 1; }
            }
            if (removed) free(f[c].l); else f[w++] = f[c];
        }
        fm = w;
        if (status == ORC_UNDEF) {
            int32_t last = 0, curn = *n_solved;
            while (curn != last && status == ORC_UNDEF) {
                /* propagate_literals */
                int32_t c = 0;
                while (c < fm) {
                    pclause cl = f[c];
                    int32_t nfalse = 0, unit = -1, sat = 0;
                    for (int32_t i = 0; i < cl.n; i++) {
                        int32_t l = cl.l[i];
                        int cur_false = 0;
                        for (int32_t j = 0; j < *n_solved; j++) {
                            if (l == solved[j]) { sat = 1; break; }
                            if (l == LIT_NEG(solved[j])) { nfalse++; cur_false = 1; break; }
                        }
                        if (sat) break;
                        if (!cur_false) unit = l;
                    }
                    if (sat) unit = -1;
                    else if (nfalse != cl.n - 1) unit = -1;
                    if (unit != -1) {
                        /* UnaryClausesRemover::add */
                        int dup = 0;
                        for (int32_t j = 0; j < *n_solved; j++) {
                            if (solved[j] == unit) { dup = 1; break; }
                            if (solved[j] == LIT_NEG(unit)) { status = ORC_UNSAT; dup = 1; break; }
                        }
                        if (!dup) solved[(*n_solved)++] = unit;
                        if (status == ORC_UNSAT) break;
                    }
                    if (sat || unit != -1) {
                        free(f[c].l);
                        memmove(f + c, f + c + 1, sizeof(pclause) * (fm - c - 1));
                        fm--;
                    } else if (!sat && nfalse == cl.n) {
                        status = ORC_UNSAT;
                        break;
                    } else {
                        /* clean_clause: drop literals falsified by solved literals */
                        int32_t k = 0;
                        for (int32_t i = 0; i < f[c].n; i++) {
                            int32_t l = f[c].l[i];
                            int fl = 0;
                            for (int32_t j = 0; j < *n_solved; j++)
                                if (l == LIT_NEG(solved[j])) { fl = 1; break; }
                            if (!fl) f[c].l[k++] = l;
                        }
                        f[c].n = k;
                        c++;
                    }
                }
                if (status == ORC_UNDEF && fm == 0) status = ORC_SAT;
                last = curn;
                curn = *n_solved;
            }
        }
    }
    int64_t pos = 0;
    *out_m = (status == ORC_UNDEF) ? fm : 0;
    out_clause_off[0] = 0;
    for (int32_t c = 0; c < fm; c++) {
        if (status == ORC_UNDEF) {
            memcpy(out_lits + pos, f[c].l, sizeof(int32_t) * f[c].n);
            pos += f[c].n;
            out_clause_off[c + 1] = pos;
        }
        free(f[c].l);
    }
    free(f);
    return status;
}

/* value of literal l given var values val[] in {-1,0,1}: 1 true, 0 false, -1 undef */
static inline int lit_value(const int8_t *val, int32_t l)
{
    int8_t x = val[LIT_VAR(l)];
    if (x < 0) return -1;
    return (l & 1) ? x : !x;
}

/* a6..a13, the reference's sequential solver, is restated in refsolve.c. */

/* ------------------------------------------------------------------------------------------
 * Independent verdict checkers (no reference counterpart). */
int orc_verify_model(const int32_t *lits, const int64_t *clause_off, int32_t m, const int8_t *model)
{
    for (int32_t c = 0; c < m; c++) {
        int ok = 0;
        for (int64_t k = clause_off[c]; k < clause_off[c + 1] && !ok; k++) {
            int32_t l = lits[k];
            int8_t x = model[LIT_VAR(l)];
            if (x >= 0 && ((l & 1) ? x : !x)) ok = 1;
        }
        if (!ok) return 0;
    }
    return 1;
}

int orc_solve_brute(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n, int8_t *model)
{
    if (n > 26) return ORC_UNDEF;
    uint64_t *pos = (uint64_t *)calloc(m ? m : 1, 8), *neg = (uint64_t *)calloc(m ? m : 1, 8);
    for (int32_t c = 0; c < m; c++)
        for (int64_t k = clause_off[c]; k < clause_off[c + 1]; k++) {
            int32_t l = lits[k];
            if (l & 1) pos[c] |= 1ull << LIT_VAR(l); else neg[c] |= 1ull << LIT_VAR(l);
        }
    int status = ORC_UNSAT;
    for (uint64_t x = 0; x < (1ull << n); x++) {
        int ok = 1;
        for (int32_t c = 0; c < m && ok; c++) if (!((x & pos[c]) | (~x & neg[c]))) ok = 0;
        if (ok) {
            status = ORC_SAT;
            if (model) for (int32_t v = 0; v < n; v++) model[v] = (int8_t)((x >> v) & 1);
            break;
        }
    }
    free(pos); free(neg);
    return status;
}

/* DPLL with full-pass unit propagation and chronological backtracking. */
typedef struct { const int32_t *lits; const int64_t *off; int32_t m, n; int8_t *val; int64_t nodes, max_nodes; int32_t *stk; } dpll_t;

static int dpll_up(dpll_t *d, int32_t *sp)
{
    int changed = 1;
    while (changed) {
        changed = 0;
        for (int32_t c = 0; c < d->m; c++) {
            int32_t nund = 0, ul = -1, sat = 0;
            for (int64_t k = d->off[c]; k < d->off[c + 1]; k++) {
                int lv = lit_value(d->val, d->lits[k]);
                if (lv == 1) { sat = 1; break; }
                if (lv < 0) { nund++; ul = d->lits[k]; }
            }
            if (sat) continue;
            if (nund == 0) return 0;
            if (nund == 1 && lit_value(d->val, ul) < 0) {
                d->val[LIT_VAR(ul)] = (int8_t)(ul & 1);
                d->stk[(*sp)++] = LIT_VAR(ul);
                changed = 1;
            }
        }
    }
    return 1;
}

static int dpll_rec(dpll_t *d, int32_t sp)
{
    if (d->max_nodes > 0 && ++d->nodes > d->max_nodes) return -1;
    int32_t sp0 = sp;
    if (!dpll_up(d, &sp)) { while (sp > sp0) d->val[d->stk[--sp]] = -1; return 0; }
    int32_t v = -1;
    for (int32_t i = 0; i < d->n; i++) if (d->val[i] < 0) { v = i; break; }
    if (v < 0) return 1;
    for (int b = 0; b < 2; b++) {
        d->val[v] = (int8_t)b;
        d->stk[sp] = v;
        int r = dpll_rec(d, sp + 1);
        if (r != 0) return r;
        d->val[v] = -1;
    }
    while (sp > sp0) d->val[d->stk[--sp]] = -1;
    return 0;
}

int orc_solve_dpll(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                   int8_t *model, int64_t max_nodes)
{
    dpll_t d = { lits, clause_off, m, n, NULL, 0, max_nodes, NULL };
    d.val = (int8_t *)malloc(n ? n : 1); memset(d.val, -1, n ? n : 1);
    d.stk = (int32_t *)malloc(sizeof(int32_t) * (n + 1) * 2);
    int r = dpll_rec(&d, 0);
    int status = r == 1 ? ORC_SAT : r == 0 ? ORC_UNSAT : ORC_UNDEF;
    if (model) for (int32_t v = 0; v < n; v++) model[v] = status == ORC_SAT ? (d.val[v] < 0 ? 0 : d.val[v]) : -1;
    free(d.val); free(d.stk);
    return status;
}

/* ------------------------------------------------------------------------------------------
 * a14: cube variables and cube enumeration. */
typedef struct { int64_t score; int32_t var; } vscore;
static int vscore_cmp(const void *a, const void *b)
{
    const vscore *x = (const vscore *)a, *y = (const vscore *)b;
    if (x->score != y->score) return x->score > y->score ? -1 : 1;
    return x->var - y->var;
}

int orc_choose_cube_vars(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                         int32_t n_dead, int64_t threads, int64_t blocks, int32_t strategy,
                         int32_t *out_vars)
{
    /* MaxClauseJobChooser::evalVarPerJobsDistribute / Uniform (JobChooser.cu:115-133) */
    int64_t working = (int64_t)n - n_dead;
    int64_t free_minus = working - 2;           /* MIN_FREE_VARS */
    if (free_minus < 1) free_minus = 1;
    int32_t d;
    if (strategy == 0) {
        double expected = (double)(threads * blocks) * 10.0;   /* JOBS_PER_THREAD */
        int64_t e = free_minus < 62 ? free_minus : 62;
        double mx = 1.0;
        for (int64_t i = 0; i < e; i++) mx *= 2.0;
        if (expected > mx) expected = mx;
        int32_t lg = 0;
        while ((double)(1ull << (lg + 1)) <= expected && lg < 62) lg++;
        d = lg + 1;
        if (d > 15) d = 15;                      /* MAX_VARS */
    } else {
        int64_t e = 7;                           /* UNIFORM_NUMBER_OF_VARS */
        if (free_minus < e) e = free_minus;
        d = (int32_t)(e < 15 ? e : 15);
    }
    if (d > n) d = n;
    /* VariableChooser::evaluate: score(v) += v per occurrence (VariableChooser.cu:23-40) */
    vscore *sc = (vscore *)malloc(sizeof(vscore) * (n ? n : 1));
    for (int32_t v = 0; v < n; v++) { sc[v].var = v; sc[v].score = 0; }
    for (int64_t k = clause_off[0]; k < clause_off[m]; k++) { int32_t v = LIT_VAR(lits[k]); sc[v].score += v; }
    qsort(sc, n, sizeof(vscore), vscore_cmp);
    for (int32_t i = 0; i < d; i++) out_vars[i] = sc[i].var;
    free(sc);
    return d;
}

void orc_enumerate_cubes(const int32_t *vars, int32_t d, int32_t *out)
{
    int64_t nc = 1ll << d;
    for (int64_t c = 0; c < nc; c++)
        for (int32_t j = 0; j < d; j++)
            out[c * d + j] = 2 * vars[j] + (((c >> (d - 1 - j)) & 1) ? 0 : 1);
}

void orc_restart_schedule(int32_t first, float factor, int32_t k, int32_t *out)
{
    int32_t t = first;
    for (int32_t i = 0; i < k; i++) { out[i] = t; t = (int32_t)(t * factor); }
}

/* ------------------------------------------------------------------------------------------
 * The reference's sequential path for a batch, one instance at a time, as its main does
 * (src/SATSolver/main.cu:125-211): preprocessing (FormulaData.cu:82-105), an answer from
 * preprocessing if it gives one (main.cu:154-163), else the sequential solver with the
 * solved literals' vars as dead vars (main.cu:165-174).  status[i] gets the verdict. */
void orc_solve_ref_batch(int32_t n_inst, const int32_t *lits, const int64_t *clause_off,
                         const int64_t *inst_clause_off, const int32_t *n_vars, int32_t max_iterations,
                         int32_t *status)
{
    for (int32_t i = 0; i < n_inst; i++) {
        int64_t c0 = inst_clause_off[i], c1 = inst_clause_off[i + 1];
        int32_t m = (int32_t)(c1 - c0), n = n_vars[i];
        int64_t L = clause_off[c1] - clause_off[c0];
        int32_t *pl = (int32_t *)malloc(sizeof(int32_t) * (L ? L : 1));
        int64_t *po = (int64_t *)malloc(sizeof(int64_t) * (m + 1));
        int64_t *io = (int64_t *)malloc(sizeof(int64_t) * (m + 1));
        int32_t *solved = (int32_t *)malloc(sizeof(int32_t) * (m + 1));
        for (int32_t c = 0; c <= m; c++) io[c] = clause_off[c0 + c] - clause_off[c0];
        int32_t pm = 0, ns = 0;
        int st = orc_preprocess(lits + clause_off[c0], io, m, n, pl, po, &pm, solved, &ns);
        if (st == ORC_UNDEF) {
            uint8_t *dead = (uint8_t *)calloc(n ? n : 1, 1);
            for (int32_t k = 0; k < ns; k++) dead[solved[k] >> 1] = 1;
            st = orc_solve_ref(pl, po, pm, n, dead, max_iterations, NULL, NULL);
            free(dead);
        }
        status[i] = st;
        free(pl); free(po); free(io); free(solved);
    }
}

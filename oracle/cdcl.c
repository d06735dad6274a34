/*
 * cdcl.c -- textbook CDCL used only as an independent verdict checker in tests
 * (TEST INFRASTRUCTURE ONLY; no reference counterpart).
 *
 * 1-UIP learning (Marques-Silva & Sakallah, GRASP 1999), VSIDS with a binary heap
 * and phase saving (Moskewicz et al., Chaff 2001; Pipatsrisawat & Darwiche 2007),
 * Luby restarts (Luby, Sinclair & Zuckerman 1993) and activity-based learned-clause
 * deletion (Een & Sorensson, MiniSat 2003).  It shares no code with the reference-
 * scheme solver in oracle.c, so a verdict on which both agree is checked twice.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define VAR(l) ((l) >> 1)
#define NEG(l) ((l) ^ 1)

typedef struct { int32_t *d; int32_t n, cap; } vec;
static inline void vpush(vec *v, int32_t x)
{
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 4; v->d = (int32_t *)realloc(v->d, sizeof(int32_t) * v->cap); }
    v->d[v->n++] = x;
}

typedef struct {
    int32_t n;
    /* clauses */
    int32_t *lits; int64_t nl, capl;
    int64_t *cs; int32_t *cn; uint8_t *learnt, *dead; double *cact; int32_t nc, capc;
    vec *watch;
    int8_t *val; int8_t *phase; int32_t *lvl, *rsn; uint8_t *seen;
    int32_t *trail, ntrail, qhead, *lim, nlim;
    double *act, inc, cinc;
    int32_t *heap, hn, *hpos;
} cs_t;

static inline int val_lit(const cs_t *s, int32_t l)
{
    int8_t x = s->val[VAR(l)];
    return x < 0 ? -1 : ((l & 1) ? x : !x);
}

/* binary max-heap on activity */
static void h_up(cs_t *s, int32_t i)
{
    int32_t v = s->heap[i];
    while (i > 0) {
        int32_t p = (i - 1) >> 1;
        if (s->act[s->heap[p]] >= s->act[v]) break;
        s->heap[i] = s->heap[p]; s->hpos[s->heap[i]] = i; i = p;
    }
    s->heap[i] = v; s->hpos[v] = i;
}
static void h_down(cs_t *s, int32_t i)
{
    int32_t v = s->heap[i];
    for (;;) {
        int32_t c = 2 * i + 1;
        if (c >= s->hn) break;
        if (c + 1 < s->hn && s->act[s->heap[c + 1]] > s->act[s->heap[c]]) c++;
        if (s->act[s->heap[c]] <= s->act[v]) break;
        s->heap[i] = s->heap[c]; s->hpos[s->heap[i]] = i; i = c;
    }
    s->heap[i] = v; s->hpos[v] = i;
}
static void h_insert(cs_t *s, int32_t v)
{
    if (s->hpos[v] >= 0) return;
    s->heap[s->hn] = v; s->hpos[v] = s->hn++; h_up(s, s->hn - 1);
}
static int32_t h_pop(cs_t *s)
{
    int32_t v = s->heap[0];
    s->hpos[v] = -1;
    if (--s->hn > 0) { s->heap[0] = s->heap[s->hn]; s->hpos[s->heap[0]] = 0; h_down(s, 0); }
    return v;
}

static int32_t add_clause(cs_t *s, const int32_t *l, int32_t n, int learnt)
{
    if (s->nl + n > s->capl) { while (s->nl + n > s->capl) s->capl = 2 * s->capl + 16; s->lits = (int32_t *)realloc(s->lits, 4 * s->capl); }
    if (s->nc + 1 > s->capc) {
        s->capc = 2 * s->capc + 16;
        s->cs = (int64_t *)realloc(s->cs, 8 * s->capc); s->cn = (int32_t *)realloc(s->cn, 4 * s->capc);
        s->learnt = (uint8_t *)realloc(s->learnt, s->capc); s->dead = (uint8_t *)realloc(s->dead, s->capc);
        s->cact = (double *)realloc(s->cact, 8 * s->capc);
    }
    memcpy(s->lits + s->nl, l, 4 * n);
    int32_t c = s->nc++;
    s->cs[c] = s->nl; s->cn[c] = n; s->learnt[c] = (uint8_t)learnt; s->dead[c] = 0; s->cact[c] = 0;
    s->nl += n;
    if (n >= 2) { vpush(&s->watch[NEG(l[0])], c); vpush(&s->watch[NEG(l[1])], c); }
    return c;
}

static void assign(cs_t *s, int32_t l, int32_t r)
{
    int32_t v = VAR(l);
    s->val[v] = (int8_t)(l & 1); s->lvl[v] = s->nlim; s->rsn[v] = r;
    s->trail[s->ntrail++] = l;
}

/* watches are indexed by the negation of the watched literal: watch[p] holds clauses
 * whose first or second literal is NEG(p), visited when p becomes true */
static int32_t propagate(cs_t *s)
{
    while (s->qhead < s->ntrail) {
        int32_t p = s->trail[s->qhead++];
        int32_t fl = NEG(p);
        vec *ws = &s->watch[p];
        int32_t i = 0, j = 0;
        while (i < ws->n) {
            int32_t c = ws->d[i++];
            if (s->dead[c]) continue;
            int32_t *cl = s->lits + s->cs[c];
            if (cl[0] == fl) { cl[0] = cl[1]; cl[1] = fl; }
            if (val_lit(s, cl[0]) == 1) { ws->d[j++] = c; continue; }
            int found = 0;
            for (int32_t k = 2; k < s->cn[c]; k++)
                if (val_lit(s, cl[k]) != 0) {
                    cl[1] = cl[k]; cl[k] = fl;
                    vpush(&s->watch[NEG(cl[1])], c);
                    found = 1; break;
                }
            if (found) continue;
            ws->d[j++] = c;
            if (val_lit(s, cl[0]) == 0) {
                while (i < ws->n) ws->d[j++] = ws->d[i++];
                ws->n = j;
                return c;
            }
            assign(s, cl[0], c);
        }
        ws->n = j;
    }
    return -1;
}

static void bump_var(cs_t *s, int32_t v)
{
    if ((s->act[v] += s->inc) > 1e100) {
        for (int32_t i = 0; i < s->n; i++) s->act[i] *= 1e-100;
        s->inc *= 1e-100;
    }
    if (s->hpos[v] >= 0) h_up(s, s->hpos[v]);
}

static void cancel_until(cs_t *s, int32_t level)
{
    if (s->nlim <= level) return;
    for (int32_t i = s->ntrail - 1; i >= s->lim[level]; i--) {
        int32_t v = VAR(s->trail[i]);
        s->phase[v] = s->val[v];
        s->val[v] = -1; s->rsn[v] = -1;
        h_insert(s, v);
    }
    s->ntrail = s->qhead = s->lim[level];
    s->nlim = level;
}

static double luby(double y, int64_t x)
{
    int64_t size = 1, seq = 0;
    while (size < x + 1) { seq++; size = 2 * size + 1; }
    while (size - 1 != x) { size = (size - 1) >> 1; seq--; x = x % size; }
    double r = 1; for (int64_t i = 0; i < seq; i++) r *= y;
    return r;
}

static cs_t *g_sort_ctx;
static int cmp_cls(const void *a, const void *b)
{
    double x = g_sort_ctx->cact[*(const int32_t *)a], y = g_sort_ctx->cact[*(const int32_t *)b];
    return x < y ? -1 : x > y ? 1 : 0;
}

int orc_solve_cdcl(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                   int8_t *model, int64_t max_conflicts, int64_t *stats)
{
    cs_t S, *s = &S;
    memset(s, 0, sizeof(S));
    int32_t nv = n > 0 ? n : 1;
    s->n = n;
    s->watch = (vec *)calloc(2 * nv, sizeof(vec));
    s->val = (int8_t *)malloc(nv); memset(s->val, -1, nv);
    s->phase = (int8_t *)calloc(nv, 1);
    s->lvl = (int32_t *)calloc(nv, 4); s->rsn = (int32_t *)malloc(4 * nv);
    s->seen = (uint8_t *)calloc(nv, 1);
    s->trail = (int32_t *)malloc(4 * nv); s->lim = (int32_t *)malloc(4 * (nv + 1));
    s->act = (double *)calloc(nv, 8); s->inc = 1.0; s->cinc = 1.0;
    s->heap = (int32_t *)malloc(4 * nv); s->hpos = (int32_t *)malloc(4 * nv);
    for (int32_t v = 0; v < n; v++) { s->rsn[v] = -1; s->hpos[v] = -1; }
    for (int32_t v = 0; v < n; v++) h_insert(s, v);
    int status = ORC_UNDEF;
    int64_t conflicts = 0, decisions = 0, nlearnt = 0;
    int32_t live = 0;   /* learned clauses not deleted */
    int32_t *lbuf = (int32_t *)malloc(4 * (2 * nv + 2));
    for (int32_t c = 0; c < m && status == ORC_UNDEF; c++) {
        int32_t len = (int32_t)(clause_off[c + 1] - clause_off[c]);
        const int32_t *cl = lits + clause_off[c];
        /* drop duplicate literals and tautologies */
        int32_t k = 0, taut = 0;
        for (int32_t i = 0; i < len; i++) {
            int dup = 0;
            for (int32_t j = 0; j < k; j++) { if (lbuf[j] == cl[i]) dup = 1; if (lbuf[j] == NEG(cl[i])) taut = 1; }
            if (!dup) lbuf[k++] = cl[i];
        }
        if (taut) continue;
        if (k == 0) { status = ORC_UNSAT; break; }
        if (k == 1) {
            int vl = val_lit(s, lbuf[0]);
            if (vl == 0) status = ORC_UNSAT; else if (vl < 0) assign(s, lbuf[0], -1);
            continue;
        }
        add_clause(s, lbuf, k, 0);
    }
    if (status == ORC_UNDEF && propagate(s) >= 0) status = ORC_UNSAT;
    int64_t restart_i = 0, next_restart = 100;
    int64_t conf_since = 0;
    int32_t max_learnt = m / 3 + 100;
    while (status == ORC_UNDEF) {
        int32_t confl = propagate(s);
        if (confl >= 0) {
            conflicts++; conf_since++;
            if (s->nlim == 0) { status = ORC_UNSAT; break; }
            if (max_conflicts >= 0 && conflicts > max_conflicts) break;
            /* 1-UIP */
            int32_t nl = 1, pathc = 0, p = -1, idx = s->ntrail - 1, c = confl;
            do {
                s->cact[c] += s->cinc;
                int32_t *cl = s->lits + s->cs[c];
                for (int32_t k = (p < 0) ? 0 : 1; k < s->cn[c]; k++) {
                    int32_t q = cl[k], v = VAR(q);
                    if (s->seen[v] || s->lvl[v] == 0) continue;
                    s->seen[v] = 1;
                    bump_var(s, v);
                    if (s->lvl[v] >= s->nlim) pathc++;
                    else lbuf[nl++] = q;
                }
                while (!s->seen[VAR(s->trail[idx])]) idx--;
                p = s->trail[idx--];
                c = s->rsn[VAR(p)];
                s->seen[VAR(p)] = 0;
                pathc--;
                if (pathc > 0 && c >= 0) {
                    /* reason clauses keep the implied literal first */
                    int32_t *rc = s->lits + s->cs[c];
                    if (rc[0] != p) { for (int32_t k = 1; k < s->cn[c]; k++) if (rc[k] == p) { rc[k] = rc[0]; rc[0] = p; break; } }
                }
            } while (pathc > 0);
            lbuf[0] = NEG(p);
            for (int32_t k = 1; k < nl; k++) s->seen[VAR(lbuf[k])] = 0;
            int32_t bt = 0, bi = 1;
            for (int32_t k = 1; k < nl; k++) if (s->lvl[VAR(lbuf[k])] > bt) { bt = s->lvl[VAR(lbuf[k])]; bi = k; }
            if (nl > 1) { int32_t t = lbuf[1]; lbuf[1] = lbuf[bi]; lbuf[bi] = t; }
            cancel_until(s, bt);
            if (nl == 1) assign(s, lbuf[0], -1);
            else {
                int32_t cid = add_clause(s, lbuf, nl, 1);
                s->cact[cid] = s->cinc;
                nlearnt++;
                live++;
                assign(s, lbuf[0], cid);
            }
            s->inc /= 0.95; s->cinc /= 0.999;
        } else {
            if (conf_since >= next_restart) {
                restart_i++;
                next_restart = (int64_t)(100 * luby(2.0, restart_i));
                conf_since = 0;
                cancel_until(s, 0);
            }
            /* reduce learned DB */
            if (live > max_learnt) {
                int32_t *ids = (int32_t *)malloc(4 * live), k = 0;
                for (int32_t c = 0; c < s->nc; c++) if (s->learnt[c] && !s->dead[c]) ids[k++] = c;
                g_sort_ctx = s;
                qsort(ids, k, 4, cmp_cls);
                for (int32_t i = 0; i < k / 2; i++) {
                    int32_t c = ids[i];
                    int32_t l0 = s->lits[s->cs[c]];
                    if (s->rsn[VAR(l0)] == c && val_lit(s, l0) == 1) continue;   /* locked */
                    if (s->cn[c] > 2) { s->dead[c] = 1; live--; }
                }
                free(ids);
                max_learnt += max_learnt / 10;
            }
            int32_t v = -1;
            while (s->hn > 0) { int32_t x = h_pop(s); if (s->val[x] < 0) { v = x; break; } }
            if (v < 0) { status = ORC_SAT; break; }
            decisions++;
            s->lim[s->nlim++] = s->ntrail;
            assign(s, 2 * v + (s->phase[v] ? 1 : 0), -1);
        }
    }
    if (model) for (int32_t v = 0; v < n; v++) model[v] = status == ORC_SAT ? s->val[v] : -1;
    if (stats) { stats[0] = decisions; stats[1] = conflicts; stats[2] = nlearnt; stats[3] = 0; stats[4] = restart_i; }
    for (int32_t l = 0; l < 2 * nv; l++) free(s->watch[l].d);
    free(s->watch); free(s->val); free(s->phase); free(s->lvl); free(s->rsn); free(s->seen);
    free(s->trail); free(s->lim); free(s->act); free(s->heap); free(s->hpos); free(lbuf);
    free(s->lits); free(s->cs); free(s->cn); free(s->learnt); free(s->dead); free(s->cact);
    return status;
}

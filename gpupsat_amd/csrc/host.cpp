// host.cpp -- host-side parts of the drop-in boundary: DIMACS input, the reference's
// preprocessing, and cube generation (JobChooser / JobsQueue), with their C ABI.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/gpsat.h"
#include "gpsat_internal.hpp"
#include "jobs.hpp"

namespace gpsat {

// ---------------------------------------------------------------------------------------
// JobsQueue / JobChooser
void JobsQueue::add(const int32_t* lits, size_t n)
{
    if (closed_) return;   // JobsQueue::add refuses jobs after close (JobsQueue.cu:45-49)
    lits_.insert(lits_.end(), lits, lits + n);
    off_.push_back((int32_t)lits_.size());
    largest_ = std::max(largest_, n);
}

void JobChooser::add_jobs(JobsQueue& queue)
{
    const size_t d = chosen_.size();
    std::vector<int32_t> cube(d);
    // cube c fixes var j positive iff bit (d-1-j) of c is 0: depth-first, positive first
    for (size_t c = 0; c < ((size_t)1 << d); c++) {
        for (size_t j = 0; j < d; j++) cube[j] = 2 * chosen_[j] + (((c >> (d - 1 - j)) & 1) ? 0 : 1);
        queue.add(cube.data(), d);
    }
}

MaxClauseJobChooser::MaxClauseJobChooser(const int32_t* lits, const int64_t* clause_off, int32_t m, int32_t n_vars,
                                         int32_t n_dead, int64_t threads, int64_t blocks, ChoosingStrategy strategy)
    : lits_(lits), clause_off_(clause_off), m_(m), n_vars_(n_vars), n_working_vars_((int64_t)n_vars - n_dead),
      threads_(threads), blocks_(blocks), strategy_(strategy)
{
}

void MaxClauseJobChooser::evaluate()
{
    // vars per job: at least two vars stay free (MIN_FREE_VARS), at most MAX_VARS = 15
    const int64_t spare = std::max<int64_t>(n_working_vars_ - 2, 1);
    if (strategy_ == ChoosingStrategy::DISTRIBUTE_JOBS_PER_THREAD) {
        double expected = (double)(threads_ * blocks_) * 10.0;          // JOBS_PER_THREAD
        expected = std::min(expected, std::ldexp(1.0, (int)std::min<int64_t>(spare, 62)));
        int32_t lg = 0;
        while (lg < 62 && std::ldexp(1.0, lg + 1) <= expected) lg++;
        vars_per_job_ = std::min(lg + 1, 15);
    } else {
        vars_per_job_ = (int32_t)std::min<int64_t>(std::min<int64_t>(7, spare), 15);
    }
    vars_per_job_ = std::min(vars_per_job_, n_vars_);
    // score(v) = sum of v over the occurrences of v
    std::vector<int64_t> score(n_vars_, 0);
    for (int64_t k = clause_off_[0]; k < clause_off_[m_]; k++) {
        const int32_t v = lits_[k] >> 1;
        score[v] += v;
    }
    std::vector<int32_t> order(n_vars_);
    for (int32_t v = 0; v < n_vars_; v++) order[v] = v;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return score[a] > score[b]; });
    chosen_.assign(order.begin(), order.begin() + vars_per_job_);
    evaluated_ = true;
}

void SimpleJobChooser::evaluate()
{
    std::vector<char> dead(n_vars_, 0);
    for (int32_t v : dead_)
        if (v >= 0 && v < n_vars_) dead[v] = 1;
    const int32_t live = n_vars_ - (int32_t)std::count(dead.begin(), dead.end(), 1);
    const int32_t want = std::min(live, 7);
    chosen_.clear();
    for (int32_t v = 0; v < n_vars_ && (int32_t)chosen_.size() < want; v++)
        if (!dead[v]) chosen_.push_back(v);
    evaluated_ = true;
}

void ScoredJobChooser::evaluate()
{
    std::vector<char> dead(n_vars_, 0);
    for (int32_t v : dead_)
        if (v >= 0 && v < n_vars_) dead[v] = 1;
    std::vector<int32_t> order;
    for (int32_t v = 0; v < n_vars_; v++)
        if (!dead[v]) order.push_back(v);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return scores_[a] > scores_[b]; });
    chosen_.assign(order.begin(), order.begin() + std::min<size_t>(order.size(), (size_t)std::max(depth_, 0)));
    evaluated_ = true;
}

// ---------------------------------------------------------------------------------------
// Preprocessing.  RepeatedLiteralsRemover (RepeatedLiteralsRemover.cu:25-62): the first
// occurrence of each literal stays, a clause with both polarities of a var goes; no
// clause left means SAT.  UnaryClausesRemover (UnaryClausesRemover.cu:13-173): unit
// clauses become solved literals in order (an already solved literal leaves its clause
// in place; the complement of a solved literal means UNSAT); then ordered passes over the
// formula, repeated while the solved set grows: a clause with a solved literal is
// dropped, a clause left with one non-false literal solves it and is dropped, a clause
// with every literal false means UNSAT, any other clause loses its false literals.
Preprocessed preprocess(const int32_t* lits, const int64_t* clause_off, int32_t m, int32_t n_vars)
{
    Preprocessed out;
    std::vector<std::vector<int32_t>> f;
    f.reserve(m);
    std::vector<int8_t> seen(std::max(n_vars, 1), 0);
    for (int32_t c = 0; c < m; c++) {
        std::vector<int32_t> cl;
        bool taut = false;
        for (int64_t k = clause_off[c]; k < clause_off[c + 1]; k++) {
            const int32_t l = lits[k];
            const int8_t pol = (l & 1) ? 1 : 2;
            int8_t& s = seen[l >> 1];
            if (s == 0) {
                s = pol;
                cl.push_back(l);
            } else if (s != pol) {
                taut = true;
                break;
            }
        }
        for (int64_t k = clause_off[c]; k < clause_off[c + 1]; k++) seen[lits[k] >> 1] = 0;
        if (!taut) f.push_back(std::move(cl));
    }
    if (f.empty()) {
        out.status = 0;
        return out;
    }
    // value of every literal under the solved set: 1 true, -1 false, 0 free
    std::vector<int8_t> val(2 * (size_t)std::max(n_vars, 1), 0);
    auto solve = [&](int32_t l) -> bool {   // false: already solved (either polarity)
        if (val[l] != 0) {
            if (val[l] < 0) out.status = 1;
            return false;
        }
        val[l] = 1;
        val[l ^ 1] = -1;
        out.solved.push_back(l);
        return true;
    };
    {
        std::vector<std::vector<int32_t>> g;
        g.reserve(f.size());
        for (auto& cl : f) {
            if (cl.size() == 1 && solve(cl[0])) continue;
            g.push_back(std::move(cl));
        }
        f.swap(g);
    }
    if (out.status == 2) {
        size_t last = 0, cur = out.solved.size();
        while (cur != last && out.status == 2) {
            std::vector<std::vector<int32_t>> g;
            g.reserve(f.size());
            for (size_t i = 0; i < f.size(); i++) {
                std::vector<int32_t>& cl = f[i];
                if (out.status != 2) {
                    g.push_back(std::move(cl));
                    continue;
                }
                bool sat = false;
                size_t nfalse = 0;
                int32_t unit = -1;
                for (int32_t l : cl) {
                    if (val[l] > 0) {
                        sat = true;
                        break;
                    }
                    if (val[l] < 0) nfalse++;
                    else unit = l;
                }
                if (sat) continue;
                if (nfalse + 1 == cl.size()) {
                    solve(unit);
                    if (out.status != 2) {
                        g.push_back(std::move(cl));   // the formula is not used further
                        continue;
                    }
                    continue;
                }
                if (nfalse == cl.size()) {
                    out.status = 1;
                    g.push_back(std::move(cl));
                    continue;
                }
                cl.erase(std::remove_if(cl.begin(), cl.end(), [&](int32_t l) { return val[l] < 0; }), cl.end());
                g.push_back(std::move(cl));
            }
            f.swap(g);
            if (out.status == 2 && f.empty()) out.status = 0;
            last = cur;
            cur = out.solved.size();
        }
    }
    if (out.status == 2) {
        for (auto& cl : f) {
            out.lits.insert(out.lits.end(), cl.begin(), cl.end());
            out.clause_off.push_back((int64_t)out.lits.size());
        }
    }
    return out;
}

}  // namespace gpsat

using namespace gpsat;

struct gpsat_formula {
    int32_t n_vars = 0, header_vars = 0, header_clauses = 0;
    std::vector<int32_t> lits;
    std::vector<int64_t> off{0};
};

namespace {

// DIMACS CNF (src/FileManager/CnfReader.cpp:30-135): comment lines start with 'c', one
// "p cnf V C" header, clauses end with 0.  n_vars is the highest var seen.
int parse_dimacs(const char* s, size_t len, gpsat_formula* f)
{
    size_t i = 0;
    bool header = false;
    std::vector<int32_t> cur;
    int32_t maxv = 0;
    while (i < len) {
        const char ch = s[i];
        if (ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n') {
            i++;
            continue;
        }
        if (ch == 'c' || (ch == 'p' && header)) {
            while (i < len && s[i] != '\n') i++;
            continue;
        }
        if (ch == '%') break;   // SATLIB end marker
        if (ch == 'p') {
            long v = 0, c = 0;
            char kind[8] = {0};
            std::string line;
            while (i < len && s[i] != '\n') line.push_back(s[i++]);
            if (sscanf(line.c_str(), "p %7s %ld %ld", kind, &v, &c) != 3 || strcmp(kind, "cnf") != 0 || v < 0 || c < 0)
                return GPSAT_ERR_PARSE;
            f->header_vars = (int32_t)v;
            f->header_clauses = (int32_t)c;
            header = true;
            continue;
        }
        if (!header) return GPSAT_ERR_PARSE;
        char* end = nullptr;
        const long x = strtol(s + i, &end, 10);
        if (end == s + i) return GPSAT_ERR_PARSE;
        i = (size_t)(end - s);
        if (x == 0) {
            f->lits.insert(f->lits.end(), cur.begin(), cur.end());
            f->off.push_back((int64_t)f->lits.size());
            cur.clear();
        } else {
            const long a = x < 0 ? -x : x;
            if (a > (1L << 29)) return GPSAT_ERR_PARSE;
            maxv = std::max(maxv, (int32_t)a);
            cur.push_back((int32_t)(2 * (a - 1) + (x > 0 ? 1 : 0)));
        }
    }
    if (!header) return GPSAT_ERR_PARSE;
    if (!cur.empty()) {
        f->lits.insert(f->lits.end(), cur.begin(), cur.end());
        f->off.push_back((int64_t)f->lits.size());
    }
    f->n_vars = maxv;
    return GPSAT_OK;
}

}  // namespace

extern "C" {

int gpsat_dimacs_parse(const char* text, int64_t len, gpsat_formula** out)
{
    if (!text || !out || len < 0) {
        set_error("null argument");
        return GPSAT_ERR_ARG;
    }
    gpsat_formula* f = new gpsat_formula();
    const int rc = parse_dimacs(text, (size_t)len, f);
    if (rc != GPSAT_OK) {
        delete f;
        *out = nullptr;
        set_error("Error parsing inputs.");
        return rc;
    }
    *out = f;
    return GPSAT_OK;
}

int gpsat_dimacs_read(const char* path, gpsat_formula** out)
{
    if (!path || !out) {
        set_error("null argument");
        return GPSAT_ERR_ARG;
    }
    *out = nullptr;
    FILE* fp = fopen(path, "rb");
    if (!fp) {
        set_error(std::string("cannot open ") + path);
        return GPSAT_ERR_ARG;
    }
    std::string buf;
    char tmp[1 << 16];
    size_t r;
    while ((r = fread(tmp, 1, sizeof(tmp), fp)) > 0) buf.append(tmp, r);
    fclose(fp);
    return gpsat_dimacs_parse(buf.data(), (int64_t)buf.size(), out);
}

int gpsat_formula_info(const gpsat_formula* f, int32_t* n_vars, int32_t* n_clauses, int64_t* n_lits,
                       int32_t* header_vars, int32_t* header_clauses)
{
    if (!f) {
        set_error("null formula");
        return GPSAT_ERR_ARG;
    }
    if (n_vars) *n_vars = f->n_vars;
    if (n_clauses) *n_clauses = (int32_t)(f->off.size() - 1);
    if (n_lits) *n_lits = (int64_t)f->lits.size();
    if (header_vars) *header_vars = f->header_vars;
    if (header_clauses) *header_clauses = f->header_clauses;
    return GPSAT_OK;
}

int gpsat_formula_arrays(const gpsat_formula* f, int32_t* lits, int64_t* clause_off)
{
    if (!f || (!lits && !f->lits.empty()) || !clause_off) {
        set_error("null argument");
        return GPSAT_ERR_ARG;
    }
    if (!f->lits.empty()) memcpy(lits, f->lits.data(), sizeof(int32_t) * f->lits.size());
    memcpy(clause_off, f->off.data(), sizeof(int64_t) * f->off.size());
    return GPSAT_OK;
}

int gpsat_formula_free(gpsat_formula* f)
{
    delete f;
    return GPSAT_OK;
}

int gpsat_preprocess(const int32_t* lits, const int64_t* clause_off, int32_t m, int32_t n_vars, int32_t* out_lits,
                     int64_t* out_clause_off, int32_t* out_m, int32_t* solved, int32_t* n_solved, int32_t* status)
{
    if (!clause_off || !out_clause_off || !out_m || !solved || !n_solved || !status || m < 0 || n_vars < 0) {
        set_error("null or negative argument");
        return GPSAT_ERR_ARG;
    }
    for (int64_t k = clause_off[0]; k < clause_off[m]; k++)
        if (lits[k] < 0 || (lits[k] >> 1) >= n_vars) {
            set_error("literal code out of range");
            return GPSAT_ERR_ARG;
        }
    Preprocessed p = preprocess(lits, clause_off, m, n_vars);
    *status = p.status;
    *out_m = (int32_t)(p.clause_off.size() - 1);
    if (!p.lits.empty()) memcpy(out_lits, p.lits.data(), sizeof(int32_t) * p.lits.size());
    memcpy(out_clause_off, p.clause_off.data(), sizeof(int64_t) * p.clause_off.size());
    *n_solved = (int32_t)p.solved.size();
    if (!p.solved.empty()) memcpy(solved, p.solved.data(), sizeof(int32_t) * p.solved.size());
    return GPSAT_OK;
}

int gpsat_cube_vars(const int32_t* lits, const int64_t* clause_off, int32_t m, int32_t n_vars, const int32_t* dead_vars,
                    int32_t n_dead, int64_t threads, int64_t blocks, int32_t strategy, int32_t* out_vars,
                    int32_t* depth)
{
    if (!clause_off || !out_vars || !depth || n_vars < 0 || n_dead < 0 || (n_dead && !dead_vars)) {
        set_error("null or negative argument");
        return GPSAT_ERR_ARG;
    }
    std::vector<int32_t> dead(dead_vars, dead_vars + n_dead);
    if (strategy == 0 || strategy == 1) {
        MaxClauseJobChooser ch(lits, clause_off, m, n_vars, n_dead, threads, blocks,
                               strategy == 0 ? ChoosingStrategy::DISTRIBUTE_JOBS_PER_THREAD : ChoosingStrategy::UNIFORM);
        ch.evaluate();
        *depth = (int32_t)ch.chosen_vars().size();
        std::copy(ch.chosen_vars().begin(), ch.chosen_vars().end(), out_vars);
    } else if (strategy == 2) {
        SimpleJobChooser ch(n_vars, dead);
        ch.evaluate();
        *depth = (int32_t)ch.chosen_vars().size();
        std::copy(ch.chosen_vars().begin(), ch.chosen_vars().end(), out_vars);
    } else {
        set_error("unknown cube strategy");
        return GPSAT_ERR_ARG;
    }
    return GPSAT_OK;
}

int gpsat_cube_vars_scored(const float* scores, int32_t n_vars, const int32_t* dead_vars, int32_t n_dead,
                           int32_t depth, int32_t* out_vars, int32_t* out_depth)
{
    if (!scores || !out_vars || !out_depth || depth < 0 || depth > 24 || (n_dead && !dead_vars)) {
        set_error("bad argument");
        return GPSAT_ERR_ARG;
    }
    ScoredJobChooser ch(scores, n_vars, std::vector<int32_t>(dead_vars, dead_vars + n_dead), depth);
    ch.evaluate();
    *out_depth = (int32_t)ch.chosen_vars().size();
    std::copy(ch.chosen_vars().begin(), ch.chosen_vars().end(), out_vars);
    return GPSAT_OK;
}

int gpsat_cubes(const int32_t* vars, int32_t depth, int32_t* out_lits)
{
    if (depth < 0 || depth > 24 || (depth && (!vars || !out_lits))) {
        set_error("bad cube depth or null buffer");
        return GPSAT_ERR_ARG;
    }
    struct Fixed : JobChooser {
        Fixed(const int32_t* v, int32_t d) { chosen_.assign(v, v + d); }
        void evaluate() override {}
        void getJobs(JobsQueue& q) override { add_jobs(q); }
        size_t get_n_jobs() override { return (size_t)1 << chosen_.size(); }
        bool is_evaluated() override { return true; }
    } ch(vars, depth);
    JobsQueue q;
    ch.getJobs(q);
    q.close();
    if (!q.lits().empty()) memcpy(out_lits, q.lits().data(), sizeof(int32_t) * q.lits().size());
    return GPSAT_OK;
}

}  // extern "C"

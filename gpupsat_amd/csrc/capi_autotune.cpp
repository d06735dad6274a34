// capi_autotune.cpp -- the launch-shape autotuner (no reference counterpart: the reference
// fixes its grid with -b/-t and its strategies at compile time, src/SATSolver/Configs.cuh).
//
// Solve path: short trial solves, one per candidate shape (where the job state sits,
// persistent wavefronts per CU, jobs per queue grab), each under a wall-clock deadline.  A
// candidate is first resolved by solve_plan at the caller's full job count, exactly as the
// full run would resolve it; the trial then runs that mode, wave count and chunk on a
// sample of whole instances' job runs large enough that no wavefront idles, and only shapes
// that were run can be returned.  The measure is finished jobs per millisecond of kernel
// time (propagations per millisecond when too few jobs finish to compare).  Trials never
// change verdicts: every shape runs the same search per job.
//
// Eval/BCP path: the caller's launch is timed under every shape (register-resident
// 3-literal kernel or the general one, instances per workgroup, staged clause arena) and
// the process-wide tuning keys are left at the fastest.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/gpsat.h"
#include "gpsat_internal.hpp"
#include "solver.hpp"

using namespace gpsat;

namespace {

struct Report {
    std::string text;
    void line(const char* fmt, ...) __attribute__((format(printf, 2, 3)))
    {
        char buf[256];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        text += buf;
        text += '\n';
    }
    void copy_to(char* out, int32_t len) const
    {
        if (!out || len <= 0) return;
        const size_t k = std::min<size_t>(text.size(), (size_t)len - 1);
        memcpy(out, text.data(), k);
        out[k] = 0;
    }
};

// The sampled job list: whole runs of consecutive jobs of one instance (all the cubes the
// caller queued for it, so chunked cube reuse is part of what is measured), every s-th run
// of the caller's list so the sample spreads over the queue, about `target` jobs in all --
// or all of them when they are at most twice that many: the end of a run is set by its
// slowest jobs, so a sample of half the jobs can rank shapes differently from the run itself
// (on 20,000 n=50 instances a 10,000-job sample preferred a shape 7% slower in full).
struct Sample {
    std::vector<int32_t> inst, off, lits;
};

Sample sample_jobs(const int32_t* job_inst, const int32_t* cube_off, const int32_t* cube_lits, int32_t n_jobs,
                   int32_t n_inst, int64_t target)
{
    Sample s;
    s.off.push_back(0);
    const int32_t total = job_inst ? n_jobs : n_inst;
    const int64_t stride =
        total <= 2 * target ? 1 : std::max<int64_t>(1, (total + target - 1) / std::max<int64_t>(target, 1));
    if (!job_inst) {
        for (int32_t i = 0; i < n_inst; i += (int32_t)stride) {
            s.inst.push_back(i);
            s.off.push_back(0);
        }
        s.lits.push_back(0);
        return s;
    }
    int64_t r = 0;
    for (int32_t j = 0; j < n_jobs; r++) {
        int32_t e = j + 1;
        while (e < n_jobs && job_inst[e] == job_inst[j]) e++;
        if (r % stride == 0)
            for (int32_t k = j; k < e; k++) {
                s.inst.push_back(job_inst[k]);
                for (int32_t q = cube_off[k]; q < cube_off[k + 1]; q++) s.lits.push_back(cube_lits[q]);
                s.off.push_back((int32_t)s.lits.size());
            }
        j = e;
    }
    if (s.lits.empty()) s.lits.push_back(0);
    return s;
}

struct Trial {
    SolvePlan plan;
    double kernel_ms = 0;
    int64_t props = 0, finished = 0;
    int32_t jobs = 0;
    bool ok = false;
};

}  // namespace

extern "C" {

static const char* mem_name(int mem) { return mem == 1 ? "LDS" : mem == 3 ? "split" : mem == 4 ? "HBM-all" : "HBM"; }

int gpsat_autotune_solve(gpsat_batch* b, const gpsat_solve_opts* base, const int32_t* job_inst, const int32_t* cube_off,
                         const int32_t* cube_lits, int32_t n_jobs, int32_t sample_jobs_max, double trial_ms,
                         gpsat_solve_opts* best, char* report, int32_t report_len)
{
    if (!b || !best) {
        set_error("null batch or output");
        return GPSAT_ERR_ARG;
    }
    gpsat_solve_opts o;
    if (base) o = *base;
    else gpsat_solve_opts_default(&o);
    const int32_t N = b->d.n_inst;
    *best = o;
    if (N == 0) return GPSAT_OK;
    if (job_inst && (!cube_off || n_jobs <= 0)) {
        set_error("job arrays without cube offsets");
        return GPSAT_ERR_ARG;
    }
    if (sample_jobs_max <= 0) sample_jobs_max = 4096;
    if (!(trial_ms > 0)) trial_ms = 300.0;
    const bool cubes = job_inst != nullptr;
    const int32_t n_full = cubes ? n_jobs : N;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        set_error("no HIP device");
        return GPSAT_ERR_NODEVICE;
    }
    Report rep;
    rep.line("autotune solve: %d jobs, trial deadline %.0f ms; every trial runs the shape the full run resolves to",
             n_full, trial_ms);

    // The shape a full run of all n_full jobs launches with these choices.
    auto plan_for = [&](int32_t mem, int32_t waves, int32_t chunk, SolvePlan& p) -> bool {
        gpsat_solve_opts t = o;
        t.clause_mem = mem;
        t.waves = waves;
        t.chunk = chunk;
        SolveArgs a{};
        return solve_plan(b->d, t, n_full, a, p) == GPSAT_OK;
    };
    // the candidate shapes of the first round: clause memory x wavefronts per CU
    std::vector<SolvePlan> cand;
    for (int32_t mem : {1, 2, 3})
        for (int32_t per_cu : {8, 16, 24, 32}) {
            SolvePlan p;
            if (plan_for(mem, per_cu * cus, 0, p)) cand.push_back(p);
        }
    if (cand.empty()) {
        set_error(std::string("no launch shape fits this batch: ") + gpsat_last_error());
        rep.copy_to(report, report_len);
        return GPSAT_ERR_CAPACITY;
    }
    // One sample for every trial, so their finished-job counts compare: at least two chunks
    // per wavefront of the widest candidate (no wavefront idles before the deadline).
    int64_t most = 0;
    for (const SolvePlan& p : cand) most = std::max(most, (int64_t)p.waves * std::max(p.chunk, cubes ? 64 : 1));
    const Sample s = sample_jobs(job_inst, cube_off, cube_lits, n_jobs, N, std::max<int64_t>(sample_jobs_max, 2 * most));
    const int32_t J = (int32_t)s.inst.size();
    rep.line("sample: %d jobs", J);
    std::vector<Trial> done;
    // One trial: the sample launched with the full run's mode, waves and chunk.
    auto trial = [&](const SolvePlan& p) -> Trial {
        Trial t;
        t.plan = p;
        for (const Trial& d : done)
            if (d.plan.mode == p.mode && d.plan.waves == p.waves && d.plan.chunk == p.chunk) return d;
        gpsat_solve_opts to = o;
        to.clause_mem = p.mode;
        to.waves = p.waves;
        to.chunk = p.chunk;
        to.verify = 0;
        to.epoch_conflicts = 0;      // one launch per trial; epochs only replay the same kernel
        to.deadline_ms = trial_ms;
        to.learned_out = nullptr;
        to.learned_cap = 0;
        std::vector<int32_t> status(N);
        std::vector<int64_t> stats((size_t)J * GPSAT_JOB_STATS);
        gpsat_solve_info inf;
        const int rc = gpsat_solve(b, &to, s.inst.data(), s.off.data(), s.lits.data(), J, nullptr, nullptr,
                                   status.data(), nullptr, nullptr, stats.data(), nullptr, &inf);
        if (rc != GPSAT_OK) {
            rep.line("  mem=%s waves=%d chunk=%d: failed (%s)", mem_name(p.mode), p.waves, p.chunk, gpsat_last_error());
            done.push_back(t);
            return t;
        }
        if (inf.waves != p.waves || inf.chunk != p.chunk || inf.clause_mem != p.mode) {
            rep.line("  mem=%s waves=%d chunk=%d: the trial ran waves=%d chunk=%d mem=%s instead; not ranked",
                     mem_name(p.mode), p.waves, p.chunk, inf.waves, inf.chunk, mem_name(inf.clause_mem));
            done.push_back(t);
            return t;
        }
        for (int32_t j = 0; j < J; j++) {
            t.props += stats[(size_t)j * GPSAT_JOB_STATS + 3];
            const int64_t js = stats[(size_t)j * GPSAT_JOB_STATS];
            t.finished += js == 0 || js == 1;
        }
        t.kernel_ms = inf.kernel_ms;
        t.jobs = J;
        t.ok = inf.kernel_ms > 0;
        rep.line("  mem=%s waves=%d chunk=%d: %lld finished (%.2f per ms), %.0f props/ms, %.1f ms", mem_name(p.mode),
                 p.waves, p.chunk, (long long)t.finished, t.finished / std::max(t.kernel_ms, 1e-9),
                 t.props / std::max(t.kernel_ms, 1e-9), t.kernel_ms);
        done.push_back(t);
        return t;
    };
    // Ranking: finished jobs (SAT or refuted) per ms of kernel time; when no trial of the
    // round finished 32 jobs there are too few to compare and props per ms decides.
    auto pick = [&](const std::vector<Trial>& ts) -> int {
        int64_t most = 0;
        for (const Trial& t : ts)
            if (t.ok) most = std::max(most, t.finished);
        const bool by_jobs = most >= 32;
        int bi = -1;
        double bs = -1;
        for (int i = 0; i < (int)ts.size(); i++) {
            if (!ts[i].ok) continue;
            const double sc = (by_jobs ? (double)ts[i].finished : (double)ts[i].props) / ts[i].kernel_ms;
            if (sc > bs) {
                bs = sc;
                bi = i;
            }
        }
        return bi;
    };

    {   // warm-up launch (module load, first-touch of the queue buffers); not ranked
        SolvePlan p;
        if (plan_for(0, 0, 0, p)) {
            gpsat_solve_opts to = o;
            to.verify = 0;
            to.epoch_conflicts = 0;
            to.deadline_ms = std::min(trial_ms, 20.0);
            to.learned_out = nullptr;
            to.learned_cap = 0;
            const Sample w = sample_jobs(job_inst, cube_off, cube_lits, n_jobs, N, 256);
            std::vector<int32_t> status(N);
            (void)gpsat_solve(b, &to, w.inst.data(), w.off.data(), w.lits.data(), (int32_t)w.inst.size(), nullptr,
                              nullptr, status.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
        }
    }
    // clause memory x wavefronts per CU (automatic chunk), then the chunk
    std::vector<Trial> round1;
    for (const SolvePlan& p : cand) round1.push_back(trial(p));
    int bi = pick(round1);
    if (bi < 0) {
        set_error("no launch shape fits this batch");
        rep.copy_to(report, report_len);
        return GPSAT_ERR_CAPACITY;
    }
    Trial win = round1[bi];
    if (cubes) {
        std::vector<Trial> round2{win};
        for (int32_t chunk : {1, 8, 32, 64}) {
            SolvePlan p;
            if (!plan_for(win.plan.mode, win.plan.waves, chunk, p)) continue;
            round2.push_back(trial(p));
        }
        const int b2 = pick(round2);
        if (b2 >= 0) win = round2[b2];
    }
    best->clause_mem = win.plan.mode;
    best->waves = win.plan.waves;
    best->chunk = win.plan.chunk;
    rep.line("chosen: mem=%s waves=%d chunk=%d (%lld finished jobs, %.0f props/ms in its trial)", mem_name(win.plan.mode),
             win.plan.waves, win.plan.chunk, (long long)win.finished, win.props / std::max(win.kernel_ms, 1e-9));
    rep.copy_to(report, report_len);
    return GPSAT_OK;
}

int gpsat_autotune_eval(gpsat_batch* b, const uint32_t* rows_dev, int32_t A, int32_t do_eval, int32_t do_bcp,
                        int32_t reps, char* report, int32_t report_len)
{
    if (!b || !rows_dev || A < 1 || A > 64 || !(do_eval || do_bcp)) {
        set_error("autotune_eval: null batch/rows, A outside 1..64, or nothing to time");
        return GPSAT_ERR_ARG;
    }
    DeviceBatch& d = b->d;
    if (d.n_inst == 0) return GPSAT_OK;
    if (reps <= 0) reps = 10;
    const size_t nrows = (size_t)A * d.total_words;
    void *counts = nullptr, *out = nullptr, *conf = nullptr, *props = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Report rep;
    int rc = GPSAT_OK;
    auto fail = [&](const char* what, hipError_t e) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        rc = e == hipErrorOutOfMemory ? GPSAT_ERR_NOMEM : GPSAT_ERR_HIP;
    };
    hipError_t e = hipMalloc(&counts, (size_t)d.n_inst * A * 12);
    if (e == hipSuccess) e = hipMalloc(&out, std::max<size_t>(nrows, 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&conf, (size_t)d.n_inst * A * 4);
    if (e == hipSuccess) e = hipMalloc(&props, (size_t)d.n_inst * A * 4);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) fail("autotune_eval buffers", e);
    const Tuning saved = tuning();
    Tuning best = saved;
    float best_ms = 1e30f;
    rep.line("autotune eval/BCP: %d instances, A=%d, eval=%d bcp=%d, %d launches per shape", d.n_inst, A, do_eval, do_bcp,
             reps);
    // Candidate shapes per kernel the batch can take: the 3-literal kernel (instances per
    // workgroup, the small-row threshold of its BCP), the packed general kernel (wavefronts
    // per instance, the 32-bit handoff on or off) and the CSR kernel (staged arena, instances
    // per workgroup).  Each is timed as the batch would launch it.
    static const char* const kKind[] = {"3-literal", "packed", "CSR", "CSR (HBM tables)"};
    std::vector<Tuning> cands;
    if (d.fast3)
        for (int wpb : {1, 2, 4})
            for (int s64 : {0, 8, 32}) {
                Tuning t = saved;
                t.eval_fast3 = 1;
                t.eval_waves_per_block = wpb;
                t.eval_fast3_small64 = s64;
                cands.push_back(t);
            }
    if (d.pkk)
        for (int nw : {1, 2, 3, 4})
            for (int cmax : {0, 32}) {
                Tuning t = saved;
                t.eval_fast3 = 0;
                t.eval_pkk = 1;
                t.eval_pkk_waves = nw;
                t.eval_pkk_compact = cmax;
                cands.push_back(t);
            }
    for (int stage : {0, 1})
        for (int wpb : {1, 2, 4}) {
            Tuning t = saved;
            t.eval_fast3 = 0;
            t.eval_pkk = 0;
            t.eval_stage_arena = stage;
            t.eval_waves_per_block = wpb;
            cands.push_back(t);
        }
    const auto describe = [&](const Tuning& t, char* buf, size_t len) {
        tuning() = t;
        const EvalKernel ek = eval_kernel(d);
        if (ek == EK_FAST3)
            snprintf(buf, len, "kernel=%s waves/workgroup=%d small64=%d", kKind[ek], t.eval_waves_per_block,
                     t.eval_fast3_small64);
        else if (ek == EK_PACKED)
            snprintf(buf, len, "kernel=%s waves/instance=%d handoff rows=%d", kKind[ek], t.eval_pkk_waves,
                     t.eval_pkk_compact);
        else
            snprintf(buf, len, "kernel=%s stage=%d waves/workgroup=%d", kKind[ek], t.eval_stage_arena,
                     t.eval_waves_per_block);
    };
    char what[128];
    for (const Tuning& t : cands) {
        if (rc != GPSAT_OK) break;
        tuning() = t;
        auto launch = [&]() {
            return launch_eval_bcp(d, rows_dev, A, (int32_t*)counts, (uint32_t*)out, (int32_t*)conf, (int32_t*)props,
                                   nullptr, do_eval != 0, do_bcp != 0, d.stream);
        };
        e = launch();     // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, d.stream);
        for (int r = 0; r < reps && e == hipSuccess; r++) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e1, d.stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            fail("autotune_eval launch", e);
            break;
        }
        ms /= reps;
        describe(t, what, sizeof(what));
        rep.line("  %s: %.4f ms", what, ms);
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    if (rc == GPSAT_OK) {
        describe(best, what, sizeof(what));
        rep.line("chosen: %s (%.4f ms)", what, best_ms);
    }
    tuning() = rc == GPSAT_OK ? best : saved;
    rep.copy_to(report, report_len);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(counts);
    (void)hipFree(out);
    (void)hipFree(conf);
    (void)hipFree(props);
    return rc;
}

}  // extern "C"

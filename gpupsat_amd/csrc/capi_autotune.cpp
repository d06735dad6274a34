// capi_autotune.cpp -- the launch-shape autotuner (no reference counterpart: the reference
// fixes its grid with -b/-t and its strategies at compile time, src/SATSolver/Configs.cuh).
//
// Solve path: short trial solves of a sample of the caller's jobs, one per candidate
// shape (clause part in LDS or HBM, persistent wavefronts per CU, jobs per queue grab),
// each under a wall-clock deadline; the measure is propagations per millisecond of
// kernel time, i.e. the engine's work rate on this workload.  Trials never change
// verdicts: every shape runs the same search per job.
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

// The sampled job list: up to 256 instances spread evenly over the job order, and of each
// a run of consecutive jobs (all of its jobs, or its first limit/instances cubes, so
// chunked cube reuse is part of what is measured), limit jobs in all.
struct Sample {
    std::vector<int32_t> inst, off, lits;
};

Sample sample_jobs(const int32_t* job_inst, const int32_t* cube_off, const int32_t* cube_lits, int32_t n_jobs,
                   int32_t n_inst, int32_t limit)
{
    Sample s;
    s.off.push_back(0);
    if (!job_inst) {
        // one job per instance, no cubes
        const int32_t stride = std::max(1, n_inst / std::max(1, limit));
        for (int32_t i = 0; i < n_inst && (int32_t)s.inst.size() < limit; i += stride) {
            s.inst.push_back(i);
            s.off.push_back(0);
        }
        s.lits.push_back(0);
        return s;
    }
    // runs of consecutive jobs of one instance
    std::vector<std::pair<int32_t, int32_t>> runs;
    for (int32_t j = 0; j < n_jobs;) {
        int32_t e = j + 1;
        while (e < n_jobs && job_inst[e] == job_inst[j]) e++;
        runs.emplace_back(j, e);
        j = e;
    }
    const size_t want = std::min<size_t>(runs.size(), std::min<size_t>(256, (size_t)limit));
    const size_t stride = std::max<size_t>(1, runs.size() / std::max<size_t>(1, want));
    const int32_t per_run = std::max<int32_t>(1, limit / (int32_t)std::max<size_t>(1, want));
    for (size_t r = 0; r < runs.size() && (int32_t)s.inst.size() < limit; r += stride) {
        const int32_t e = std::min(runs[r].second, runs[r].first + per_run);
        for (int32_t j = runs[r].first; j < e; j++) {
            s.inst.push_back(job_inst[j]);
            for (int32_t k = cube_off[j]; k < cube_off[j + 1]; k++) s.lits.push_back(cube_lits[k]);
            s.off.push_back((int32_t)s.lits.size());
        }
    }
    if (s.lits.empty()) s.lits.push_back(0);
    return s;
}

}  // namespace

extern "C" {

static const char* mem_name(int mem) { return mem == 1 ? "LDS" : mem == 3 ? "split" : "HBM"; }

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
    const Sample s = sample_jobs(job_inst, cube_off, cube_lits, n_jobs, N, sample_jobs_max);
    const int32_t J = (int32_t)s.inst.size();
    const bool cubes = job_inst != nullptr;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        set_error("no HIP device");
        return GPSAT_ERR_NODEVICE;
    }
    std::vector<int32_t> status(N);
    std::vector<int64_t> stats((size_t)J * GPSAT_JOB_STATS);
    Report rep;
    rep.line("autotune solve: %d sampled jobs, trial deadline %.0f ms", J, trial_ms);

    // one trial: props per ms of kernel time (-1 if the shape does not fit)
    auto trial = [&](int32_t mem, int32_t waves, int32_t chunk) -> double {
        gpsat_solve_opts t = o;
        t.clause_mem = mem;
        t.waves = waves;
        t.chunk = chunk;
        t.verify = 0;
        t.epoch_conflicts = 0;      // one launch per trial; epochs only replay the same kernel
        t.deadline_ms = trial_ms;
        gpsat_solve_info inf;
        const int rc = gpsat_solve(b, &t, s.inst.data(), s.off.data(), s.lits.data(), J, nullptr, nullptr, status.data(),
                                   nullptr, nullptr, stats.data(), nullptr, &inf);
        if (rc != GPSAT_OK) {
            rep.line("  mem=%d waves=%d chunk=%d: does not fit (%s)", mem, waves, chunk, gpsat_last_error());
            return -1.0;
        }
        int64_t props = 0, done = 0;
        for (int32_t j = 0; j < J; j++) {
            props += stats[(size_t)j * GPSAT_JOB_STATS + 3];
            done += stats[(size_t)j * GPSAT_JOB_STATS] != 2;
        }
        const double rate = inf.kernel_ms > 0 ? (double)props / inf.kernel_ms : 0.0;
        rep.line("  mem=%s waves=%d chunk=%d: %.0f props/ms, %lld/%d jobs finished, %.1f ms", mem_name(mem),
                 inf.waves, inf.chunk, rate, (long long)done, J, inf.kernel_ms);
        return rate;
    };

    // clause memory x wavefronts per CU (chunk automatic), then the chunk
    double best_rate = -1.0;
    int32_t bm = 2, bw = 0, bc = 0;
    for (int32_t mem : {1, 2, 3}) {
        for (int32_t per_cu : {8, 16, 24, 32}) {
            const double r = trial(mem, per_cu * cus, 0);
            if (r > best_rate) {
                best_rate = r;
                bm = mem;
                bw = per_cu * cus;
            }
            if (r < 0) break;   // more wavefronts of this shape do not fit either
        }
    }
    if (best_rate < 0) {
        set_error("no launch shape fits this batch");
        rep.copy_to(report, report_len);
        return GPSAT_ERR_CAPACITY;
    }
    if (cubes) {
        for (int32_t chunk : {1, 8, 32, 64}) {
            const double r = trial(bm, bw, chunk);
            if (r > best_rate) {
                best_rate = r;
                bc = chunk;
            }
        }
    }
    best->clause_mem = bm;
    best->waves = bw;
    best->chunk = bc;
    rep.line("chosen: mem=%s waves=%d chunk=%s (%.0f props/ms)", mem_name(bm), bw,
             bc ? std::to_string(bc).c_str() : "auto", best_rate);
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
    for (int fast : {1, 0}) {
        if (rc != GPSAT_OK) break;
        if (fast && !d.fast3) continue;
        for (int stage : {0, 1}) {
            if (fast && stage) continue;     // the register-resident kernel has no staged arena
            for (int wpb : {1, 2, 4}) {
                tuning().eval_fast3 = fast;
                tuning().eval_stage_arena = stage;
                tuning().eval_waves_per_block = wpb;
                auto launch = [&]() {
                    return launch_eval_bcp(d, rows_dev, A, (int32_t*)counts, (uint32_t*)out, (int32_t*)conf,
                                           (int32_t*)props, nullptr, do_eval != 0, do_bcp != 0, d.stream);
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
                rep.line("  kernel=%s stage=%d waves/workgroup=%d: %.4f ms", fast ? "3-literal" : "general", stage, wpb, ms);
                if (ms < best_ms) {
                    best_ms = ms;
                    best = tuning();
                }
            }
        }
    }
    tuning() = rc == GPSAT_OK ? best : saved;
    if (rc == GPSAT_OK)
        rep.line("chosen: kernel=%s stage=%d waves/workgroup=%d (%.4f ms)", best.eval_fast3 && d.fast3 ? "3-literal" : "general",
                 best.eval_stage_arena, best.eval_waves_per_block, best_ms);
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

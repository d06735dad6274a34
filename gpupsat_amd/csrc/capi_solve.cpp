// capi_solve.cpp -- gpsat_solve: job table set-up, persistent solve launch, GPU
// re-verification of every SAT model with the clause-evaluation kernel.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/gpsat.h"
#include "gpsat_internal.hpp"
#include "solver.hpp"

using namespace gpsat;

namespace {

struct DevMem {
    std::vector<void*> ptrs;
    ~DevMem()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T*& p, size_t count)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(q);
        p = (T*)q;
        return e;
    }
};

int hip_fail(const char* what, hipError_t e)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? GPSAT_ERR_NOMEM : GPSAT_ERR_HIP;
}

}  // namespace

#define HIP_OR(expr)                                  \
    do {                                              \
        hipError_t _e = (expr);                       \
        if (_e != hipSuccess) return hip_fail(#expr, _e); \
    } while (0)


extern "C" {

int gpsat_solve_opts_default(gpsat_solve_opts* o)
{
    if (!o) {
        set_error("null options");
        return GPSAT_ERR_ARG;
    }
    memset(o, 0, sizeof(*o));
    o->decision = 0;
    o->restart_first = 100;
    o->restart_factor = 1.3f;
    o->learnt_cap = 256;
    o->larena_cap = 0;
    o->watch_slack = 0;
    o->max_conflicts = 0;
    o->max_decisions = 0;
    o->var_decay = 0.95f;
    o->phase_init = 0;
    o->max_ms = 0.0;
    o->waves = 0;
    o->clause_mem = 0;
    o->verify = 1;
    o->chunk = 0;
    o->epoch_conflicts = 0;
    o->epoch_growth = 2.0f;
    o->max_epochs = 16;
    o->deadline_ms = 0.0;
    o->inst_conflicts = 0;
    o->job_order = 0;
    o->telemetry = 0;
    o->learned_out = nullptr;
    o->learned_cap = 0;
    return GPSAT_OK;
}

int gpsat_solve(gpsat_batch* b, const gpsat_solve_opts* opts, const int32_t* job_inst, const int32_t* cube_off,
                const int32_t* cube_lits, int32_t n_jobs, const float* act_init, const int8_t* phase_init,
                int32_t* status, uint32_t* model_rows, int32_t* verified, int64_t* job_stats, int32_t* var_bumps,
                gpsat_solve_info* info)
{
    if (!b || !status) {
        set_error("null batch or status");
        return GPSAT_ERR_ARG;
    }
    gpsat_solve_opts o;
    if (opts) o = *opts;
    else gpsat_solve_opts_default(&o);
    DeviceBatch& d = b->d;
    const int32_t N = d.n_inst;
    if (N == 0) return GPSAT_OK;
    if (o.learnt_cap < 2 || o.restart_first < 1 || !(o.restart_factor >= 1.0f) || !(o.var_decay > 0.0f) ||
        o.var_decay > 1.0f) {
        set_error("invalid solver options");
        return GPSAT_ERR_ARG;
    }
    // ---- job table
    std::vector<JobDesc> jobs;
    std::vector<int32_t> cubes;
    if (!job_inst) {
        n_jobs = N;
        jobs.resize(N);
        for (int32_t i = 0; i < N; i++) jobs[i] = JobDesc{i, 0, 0, i};
        cubes.push_back(0);
    } else {
        if (n_jobs < 0 || !cube_off || (cube_off[n_jobs] > 0 && !cube_lits)) {
            set_error("bad job arrays");
            return GPSAT_ERR_ARG;
        }
        jobs.resize(n_jobs);
        for (int32_t j = 0; j < n_jobs; j++) {
            const int32_t i = job_inst[j];
            const int32_t len = cube_off[j + 1] - cube_off[j];
            if (i < 0 || i >= N || len < 0) {
                set_error("job instance or cube out of range");
                return GPSAT_ERR_ARG;
            }
            for (int32_t k = cube_off[j]; k < cube_off[j + 1]; k++)
                if (cube_lits[k] < 0 || (cube_lits[k] >> 1) >= d.h_meta[i].n) {
                    set_error("cube literal out of range");
                    return GPSAT_ERR_ARG;
                }
            jobs[j] = JobDesc{i, cube_off[j], len, j};
        }
        cubes.assign(cube_lits ? cube_lits : nullptr, cube_lits ? cube_lits + cube_off[n_jobs] : nullptr);
        if (cubes.empty()) cubes.push_back(0);
    }
    std::vector<int32_t> ncubes(N, 0);
    for (const JobDesc& j : jobs) ncubes[j.inst]++;
    // queue order by instance size n * (L + 1): largest first lets the persistent grid end
    // on short jobs; smallest first decides the most instances before a deadline.  Ties
    // keep the caller's order, so the cubes of one instance stay adjacent.
    if (o.job_order < 0 || o.job_order > 2) {
        set_error("job_order must be 0, 1 or 2");
        return GPSAT_ERR_ARG;
    }
    if (o.job_order != 2) {
        std::vector<int64_t> cost(N);
        for (int32_t i = 0; i < N; i++)
            cost[i] = (int64_t)d.h_meta[i].n * (d.h_coff[d.h_meta[i].coff_base + d.h_meta[i].m] + 1);
        const bool desc = o.job_order == 0;
        std::stable_sort(jobs.begin(), jobs.end(), [&](const JobDesc& x, const JobDesc& y) {
            if (cost[x.inst] != cost[y.inst]) return desc ? cost[x.inst] > cost[y.inst] : cost[x.inst] < cost[y.inst];
            return x.inst < y.inst;
        });
    }
    std::vector<int64_t> var_base(N + 1, 0);
    for (int32_t i = 0; i < N; i++) var_base[i + 1] = var_base[i] + d.h_meta[i].n;
    int32_t max_L = 0;
    for (int32_t i = 0; i < N; i++) max_L = std::max(max_L, d.h_coff[d.h_meta[i].coff_base + d.h_meta[i].m]);

    // ---- per-wave layout
    SolveArgs a{};
    a.o.decision_mode = o.decision;
    a.o.restart_first = o.restart_first;
    a.o.restart_factor = o.restart_factor;
    a.o.learnt_cap = o.learnt_cap;
    a.o.watch_slack = o.watch_slack > 0 ? o.watch_slack : 8;
    a.o.larena_cap = std::max(o.larena_cap > 0 ? o.larena_cap : o.learnt_cap * 32, 2 * d.max_n + 64);
    a.o.max_conflicts = o.max_conflicts;
    a.o.max_decisions = o.max_decisions;
    a.o.var_decay = o.var_decay;
    a.o.phase_init = o.phase_init ? 1 : 0;
    int clock_khz = 0;
    if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, d.device) != hipSuccess || clock_khz <= 0)
        clock_khz = 100000;
    a.o.max_ticks = o.max_ms > 0 ? (int64_t)(o.max_ms * clock_khz) : 0;
    a.o.deadline_ticks = o.deadline_ms > 0 ? (int64_t)(o.deadline_ms * clock_khz) : 0;
    a.o.inst_conflicts = o.inst_conflicts;
    a.o.telemetry = o.telemetry ? 1 : 0;
    a.max_n = d.max_n;
    a.all3 = d.all3 ? 1 : 0;
    a.wo_cap = max_L;
    a.wl_cap = 2 * d.max_n * a.o.watch_slack;
    a.wpair_cap = 2 * (d.max_m + o.learnt_cap);
    a.lmeta_cap = 5 * o.learnt_cap;   // 4 metadata ints + 1 remap slot per clause
    const int64_t clause_ints = (int64_t)a.wo_cap + a.wl_cap + a.wpair_cap + a.lmeta_cap + a.o.larena_cap;
    a.ws_ints_per_wave = (clause_ints + 63) & ~(int64_t)63;
    const size_t state_b = solve_state_lds_bytes(d.max_n);
    const size_t lds_all = state_b + 4 * (size_t)a.ws_ints_per_wave;

    hipDeviceProp_t prop;
    HIP_OR(hipGetDeviceProperties(&prop, d.device));
    const size_t lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 65536;
    const size_t lds_block_max = prop.sharedMemPerBlock ? prop.sharedMemPerBlock : 65536;
    // clause part: 1 all in LDS, 2 all in HBM, 3 the original clauses' watch lists and
    // watch pairs in LDS, the learned clauses' storage in HBM
    const size_t lds_hyb = state_b + 4 * (size_t)(a.wo_cap + a.wpair_cap - 2 * o.learnt_cap);
    int mode;
    if (o.clause_mem == 1) mode = lds_all <= lds_block_max ? 1 : 2;
    else if (o.clause_mem == 2) mode = 2;
    else if (o.clause_mem == 3) mode = lds_hyb <= lds_block_max ? 3 : 2;
    else {
        // automatic: with few jobs the grid cannot fill the GPU and each job's latency is
        // what counts, so the clause part goes to LDS as far as it fits; with many, LDS only
        // while it leaves room for enough resident wavefronts
        const bool few_jobs = n_jobs <= 4 * prop.multiProcessorCount;
        if (lds_all <= (few_jobs ? lds_block_max : std::min<size_t>(48 * 1024, lds_block_max))) mode = 1;
        else if (few_jobs && lds_hyb <= lds_block_max) mode = 3;
        else mode = 2;
    }
    const bool use_lds = mode == 1;
    const size_t lds_bytes = mode == 1 ? lds_all : mode == 3 ? lds_hyb : state_b;
    if (lds_bytes > lds_block_max) {
        set_error("instance too large for the per-wavefront solver state in LDS");
        return GPSAT_ERR_CAPACITY;
    }
    int waves = o.waves;
    if (waves <= 0) {
        const int per_cu = (int)std::max<size_t>(
            1, std::min<size_t>(4 * kSolveWavesPerEu, lds_cu / std::max<size_t>(lds_bytes, 1)));
        waves = prop.multiProcessorCount * per_cu;
    }
    waves = std::max(1, std::min(waves, n_jobs));
    // whole workgroups (surplus wavefronts find the queue empty and exit)
    const int wpb = solve_waves_per_block(lds_bytes);
    waves = (waves + wpb - 1) / wpb * wpb;
    a.lds_per_wave = (int64_t)lds_bytes;
    // chunks of consecutive jobs let cubes of one instance reuse its watch lists; a few
    // chunks per wavefront keep the queue balanced
    a.chunk = (int32_t)std::max<int64_t>(1, std::min<int64_t>(64, (int64_t)n_jobs / ((int64_t)waves * 4)));
    if (o.chunk > 0) a.chunk = o.chunk;

    // ---- device buffers
    DevMem mem;
    JobDesc* d_jobs;
    int32_t *d_cubes, *d_head, *d_status, *d_refuted, *d_undef, *d_ncubes, *d_ws = nullptr, *d_bumps = nullptr;
    int64_t *d_vbase, *d_jstats;
    uint32_t* d_models;
    float* d_act = nullptr;
    int8_t* d_phase = nullptr;
    hipStream_t st = d.stream;
    HIP_OR(mem.alloc(d_jobs, jobs.size()));
    HIP_OR(mem.alloc(d_cubes, cubes.size()));
    HIP_OR(mem.alloc(d_head, 1));
    unsigned long long *d_tstart, *d_iconf;
    HIP_OR(mem.alloc(d_tstart, 1));
    HIP_OR(mem.alloc(d_iconf, N));
    HIP_OR(mem.alloc(d_status, N));
    HIP_OR(mem.alloc(d_refuted, N));
    HIP_OR(mem.alloc(d_undef, N));
    HIP_OR(mem.alloc(d_ncubes, N));
    HIP_OR(mem.alloc(d_vbase, N + 1));
    HIP_OR(mem.alloc(d_jstats, (size_t)std::max(n_jobs, 1) * JS_COUNT));
    HIP_OR(mem.alloc(d_models, (size_t)std::max<int64_t>(d.total_words, 1)));
    if (!use_lds) HIP_OR(mem.alloc(d_ws, (size_t)waves * a.ws_ints_per_wave));
    if (var_bumps) HIP_OR(mem.alloc(d_bumps, (size_t)var_base[N]));
    if (act_init) {
        HIP_OR(mem.alloc(d_act, (size_t)var_base[N]));
        HIP_OR(hipMemcpyAsync(d_act, act_init, sizeof(float) * var_base[N], hipMemcpyHostToDevice, st));
    }
    if (phase_init) {
        HIP_OR(mem.alloc(d_phase, (size_t)var_base[N]));
        HIP_OR(hipMemcpyAsync(d_phase, phase_init, var_base[N], hipMemcpyHostToDevice, st));
    }
    std::vector<int32_t> init_status(N, 2);
    HIP_OR(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(JobDesc) * jobs.size(), hipMemcpyHostToDevice, st));
    HIP_OR(hipMemcpyAsync(d_cubes, cubes.data(), sizeof(int32_t) * cubes.size(), hipMemcpyHostToDevice, st));
    HIP_OR(hipMemsetAsync(d_head, 0, sizeof(int32_t), st));
    HIP_OR(hipMemsetAsync(d_tstart, 0, sizeof(unsigned long long), st));
    HIP_OR(hipMemsetAsync(d_iconf, 0, sizeof(unsigned long long) * N, st));
    HIP_OR(hipMemcpyAsync(d_status, init_status.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st));
    HIP_OR(hipMemsetAsync(d_refuted, 0, sizeof(int32_t) * N, st));
    HIP_OR(hipMemsetAsync(d_undef, 0, sizeof(int32_t) * N, st));
    HIP_OR(hipMemcpyAsync(d_ncubes, ncubes.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st));
    HIP_OR(hipMemcpyAsync(d_vbase, var_base.data(), sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice, st));
    HIP_OR(hipMemsetAsync(d_models, 0, sizeof(uint32_t) * std::max<int64_t>(d.total_words, 1), st));
    HIP_OR(hipMemsetAsync(d_jstats, 0, sizeof(int64_t) * std::max(n_jobs, 1) * JS_COUNT, st));
    if (d_bumps) HIP_OR(hipMemsetAsync(d_bumps, 0, sizeof(int32_t) * var_base[N], st));

    a.meta = d.meta;
    a.lits = d.lits;
    a.coff = d.coff;
    a.jobs = d_jobs;
    a.t_start = d_tstart;
    a.cube_lits = d_cubes;
    a.n_jobs = n_jobs;
    a.n_inst = N;
    a.var_base = d_vbase;
    a.act_init = d_act;
    a.phase_init = d_phase;
    a.queue_head = d_head;
    a.inst_status = d_status;
    a.inst_refuted = d_refuted;
    a.inst_undef = d_undef;
    a.inst_conf = d_iconf;
    a.inst_ncubes = d_ncubes;
    a.model_rows = d_models;
    a.job_stats = d_jstats;
    a.var_bumps = d_bumps;
    a.ws = d_ws;
    if (o.learned_out && o.learned_cap > 0) {
        if (o.learned_cap > ((int64_t)1 << 31)) {
            set_error("learned_cap above 2^31 slots");
            return GPSAT_ERR_ARG;
        }
        HIP_OR(mem.alloc(a.learned_out, (size_t)o.learned_cap));
        HIP_OR(mem.alloc(a.learned_pos, 2));
        HIP_OR(hipMemsetAsync(a.learned_pos, 0, 2 * sizeof(unsigned long long), st));
        a.learned_need = a.learned_pos + 1;
        a.learned_cap = o.learned_cap;
    }

    const bool epochs = o.epoch_conflicts > 0;
    int32_t epochs_run = 1;
    int32_t* h_remaining = nullptr;
    if (epochs) {
        if (!(o.epoch_growth >= 1.0f) || o.max_epochs < 1 || o.max_epochs > 64) {
            set_error("invalid epoch options");
            return GPSAT_ERR_ARG;
        }
        int32_t *d_active, *d_nact, *d_epoch, *d_remaining;
        int64_t* d_budget;
        float* d_astore;
        int8_t* d_pstore;
        HIP_OR(mem.alloc(d_active, 2 * (size_t)n_jobs));
        HIP_OR(mem.alloc(d_nact, 2));
        HIP_OR(mem.alloc(d_epoch, 1));
        HIP_OR(mem.alloc(d_remaining, 1));
        HIP_OR(mem.alloc(d_budget, o.max_epochs));
        HIP_OR(mem.alloc(d_astore, (size_t)var_base[N]));
        HIP_OR(mem.alloc(d_pstore, (size_t)var_base[N]));
        std::vector<int32_t> ident(n_jobs);
        for (int32_t j = 0; j < n_jobs; j++) ident[j] = j;
        const int32_t nact[2] = {n_jobs, 0}, zero = 0;
        std::vector<int64_t> budget(o.max_epochs);
        double bud = (double)o.epoch_conflicts;
        for (int32_t e = 0; e < o.max_epochs; e++, bud *= o.epoch_growth) budget[e] = (int64_t)std::min(bud, 1e15);
        std::vector<float> astore(var_base[N], 0.0f);
        std::vector<int8_t> pstore(var_base[N], (int8_t)(o.phase_init ? 1 : 0));
        if (act_init) astore.assign(act_init, act_init + var_base[N]);
        if (phase_init) pstore.assign(phase_init, phase_init + var_base[N]);
        HIP_OR(hipMemcpyAsync(d_active, ident.data(), 4 * (size_t)n_jobs, hipMemcpyHostToDevice, st));
        HIP_OR(hipMemcpyAsync(d_nact, nact, 8, hipMemcpyHostToDevice, st));
        HIP_OR(hipMemcpyAsync(d_epoch, &zero, 4, hipMemcpyHostToDevice, st));
        HIP_OR(hipMemcpyAsync(d_budget, budget.data(), 8 * budget.size(), hipMemcpyHostToDevice, st));
        HIP_OR(hipMemcpyAsync(d_astore, astore.data(), 4 * astore.size(), hipMemcpyHostToDevice, st));
        HIP_OR(hipMemcpyAsync(d_pstore, pstore.data(), pstore.size(), hipMemcpyHostToDevice, st));
        HIP_OR(hipStreamSynchronize(st));
        HIP_OR(hipHostMalloc((void**)&h_remaining, sizeof(int32_t), hipHostMallocDefault));
        a.epoch = d_epoch;
        a.active = d_active;
        a.n_active = d_nact;
        a.epoch_budget = d_budget;
        a.act_store = d_astore;
        a.phase_store = d_pstore;
        a.act_init = d_astore;
        a.phase_init = d_pstore;
    }

    hipEvent_t e0, e1, e2;
    HIP_OR(hipEventCreate(&e0));
    HIP_OR(hipEventCreate(&e1));
    HIP_OR(hipEventCreate(&e2));
    HIP_OR(hipEventRecord(e0, st));
    if (!epochs) {
        HIP_OR(launch_solve(a, mode, waves, lds_bytes, st));
    } else {
        // one restart epoch = one replay of this graph: empty the queue, run the persistent
        // solve over the open jobs, re-queue the stopped ones, publish how many remain
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int32_t* d_remaining_ptr = nullptr;
        HIP_OR(hipMalloc(&d_remaining_ptr, sizeof(int32_t)));
        mem.ptrs.push_back(d_remaining_ptr);
        HIP_OR(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t ce = launch_epoch_begin(a, st);
        if (ce == hipSuccess) ce = launch_solve(a, mode, waves, lds_bytes, st);
        if (ce == hipSuccess) ce = launch_epoch_end(a, d_remaining_ptr, st);
        if (ce == hipSuccess) ce = hipMemcpyAsync(h_remaining, d_remaining_ptr, sizeof(int32_t), hipMemcpyDeviceToHost, st);
        hipError_t ee = hipStreamEndCapture(st, &graph);
        if (ce != hipSuccess || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipHostFree(h_remaining);
            return hip_fail("capturing the epoch graph", ce != hipSuccess ? ce : ee);
        }
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(graph);
            (void)hipHostFree(h_remaining);
            return hip_fail("hipGraphInstantiate", ie);
        }
        epochs_run = 0;
        hipError_t le = hipSuccess;
        const auto t_host = std::chrono::steady_clock::now();
        for (int32_t e = 0; e < o.max_epochs; e++) {
            le = hipGraphLaunch(exec, st);
            if (le == hipSuccess) le = hipStreamSynchronize(st);
            if (le != hipSuccess) break;
            epochs_run++;
            if (*h_remaining == 0) break;
            // past the deadline every further epoch would only mark its jobs UNDEF
            const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host).count();
            if (o.deadline_ms > 0 && el > o.deadline_ms) break;
        }
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        (void)hipHostFree(h_remaining);
        if (le != hipSuccess) return hip_fail("epoch replay", le);
    }
    HIP_OR(hipEventRecord(e1, st));
    // re-verify every model: clause counts of the model rows (one row per instance)
    int32_t* d_counts = nullptr;
    if (o.verify) {
        HIP_OR(mem.alloc(d_counts, (size_t)N * 3));
        HIP_OR(launch_eval_bcp(d, d_models, 1, d_counts, nullptr, nullptr, nullptr, nullptr, true, false, st));
    }
    HIP_OR(hipEventRecord(e2, st));
    HIP_OR(hipStreamSynchronize(st));
    float ms_solve = 0, ms_verify = 0;
    (void)hipEventElapsedTime(&ms_solve, e0, e1);
    (void)hipEventElapsedTime(&ms_verify, e1, e2);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);

    HIP_OR(hipMemcpy(status, d_status, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    std::vector<int32_t> counts;
    if (o.verify) {
        counts.resize((size_t)N * 3);
        HIP_OR(hipMemcpy(counts.data(), d_counts, sizeof(int32_t) * N * 3, hipMemcpyDeviceToHost));
    }
    if (model_rows)
        HIP_OR(hipMemcpy(model_rows, d_models, sizeof(uint32_t) * d.total_words, hipMemcpyDeviceToHost));
    std::vector<int64_t> js((size_t)std::max(n_jobs, 1) * JS_COUNT);
    HIP_OR(hipMemcpy(js.data(), d_jstats, sizeof(int64_t) * js.size(), hipMemcpyDeviceToHost));
    static_assert(JS_COUNT == GPSAT_JOB_STATS, "job_stats stride of the C ABI");
    for (int32_t j = 0; j < n_jobs; j++)   // device clock ticks -> nanoseconds
        for (int f = JS_TICKS; f < JS_COUNT; f++) {
            int64_t& t = js[(size_t)j * JS_COUNT + f];
            t = (int64_t)((double)t * 1e6 / (double)clock_khz);
        }
    if (job_stats) memcpy(job_stats, js.data(), sizeof(int64_t) * (size_t)n_jobs * JS_COUNT);
    if (var_bumps) HIP_OR(hipMemcpy(var_bumps, d_bumps, sizeof(int32_t) * var_base[N], hipMemcpyDeviceToHost));

    gpsat_solve_info inf{};
    inf.kernel_ms = ms_solve;
    inf.verify_ms = ms_verify;
    inf.waves = waves;
    inf.clause_mem = mode;
    inf.lds_bytes = (int64_t)lds_bytes;
    inf.workspace_bytes = use_lds ? 0 : (int64_t)waves * a.ws_ints_per_wave * 4;
    inf.epochs = epochs_run;
    inf.chunk = a.chunk;
    if (a.learned_out) {
        unsigned long long pos[2];
        HIP_OR(hipMemcpy(pos, a.learned_pos, sizeof(pos), hipMemcpyDeviceToHost));
        inf.learned_len = (int64_t)pos[0];
        inf.learned_need = (int64_t)pos[1];
        if (pos[0]) HIP_OR(hipMemcpy(o.learned_out, a.learned_out, sizeof(int32_t) * pos[0], hipMemcpyDeviceToHost));
    }
    for (int32_t j = 0; j < n_jobs; j++)
        if (js[(size_t)j * JS_COUNT + JS_STATUS] == 4) inf.n_errors++;
    for (int32_t i = 0; i < N; i++) {
        if (status[i] == 0) {
            inf.n_sat++;
            const bool ok = o.verify && counts[(size_t)i * 3] == d.h_meta[i].m;
            if (ok) inf.n_verified++;
            if (verified) verified[i] = ok ? 1 : 0;
        } else {
            if (status[i] == 1) inf.n_unsat++;
            else inf.n_undef++;
            if (verified) verified[i] = 0;
        }
    }
    if (info) *info = inf;
    if (inf.n_errors) {
        set_error("solver reported an internal error on some jobs");
        return GPSAT_ERR_HIP;
    }
    if (o.verify && inf.n_verified != inf.n_sat) {
        set_error("a SAT model failed GPU re-verification");
        return GPSAT_ERR_HIP;
    }
    return GPSAT_OK;
}

}  // extern "C"


extern "C" int gpsat_solve_ref(gpsat_batch* b, const uint8_t* dead, int32_t max_iterations, int32_t learned_cap,
                               int32_t learned_lits, int32_t* status, uint32_t* model_rows, int64_t* stats)
{
    if (!b || !status) {
        set_error("null batch or status");
        return GPSAT_ERR_ARG;
    }
    if (learned_cap < 0 || learned_lits < 0) {
        set_error("negative capacity");
        return GPSAT_ERR_ARG;
    }
    if (learned_cap == 0) learned_cap = 4096;
    if (learned_lits == 0) learned_lits = 65536;
    DeviceBatch& d = b->d;
    const int32_t N = d.n_inst;
    if (N == 0) return GPSAT_OK;
    std::vector<RefMeta> rm(N);
    std::vector<int64_t> var_base(N + 1, 0);
    int64_t ws = 0;
    for (int32_t i = 0; i < N; i++) {
        const InstMeta& mi = d.h_meta[i];
        const int32_t L = d.h_coff[mi.coff_base + mi.m];
        if ((int64_t)mi.m + learned_cap > INT32_MAX / 2 || (int64_t)L + learned_lits > INT32_MAX / 2) {
            set_error("instance too large for the trajectory mode");
            return GPSAT_ERR_CAPACITY;
        }
        rm[i].ws = ws;
        rm[i].ccap = mi.m + learned_cap;
        rm[i].lcap = L + learned_lits;
        rm[i].wtot = solve_ref_wtot(mi.n, L);
        rm[i].pad = 0;
        ws += (solve_ref_ws_ints(mi.n, mi.m, L, rm[i].ccap, rm[i].lcap) + 15) & ~(int64_t)15;
        var_base[i + 1] = var_base[i] + mi.n;
    }
    hipStream_t st = d.stream;
    DevMem mem;
    RefMeta* d_rm;
    int32_t *d_ws, *d_status;
    int64_t *d_vb, *d_stats;
    uint32_t* d_models;
    uint8_t* d_dead = nullptr;
    HIP_OR(mem.alloc(d_rm, N));
    HIP_OR(mem.alloc(d_ws, (size_t)ws));
    HIP_OR(mem.alloc(d_status, N));
    HIP_OR(mem.alloc(d_vb, N + 1));
    HIP_OR(mem.alloc(d_stats, (size_t)N * 5));
    HIP_OR(mem.alloc(d_models, (size_t)std::max<int64_t>(d.total_words, 1)));
    if (dead) {
        HIP_OR(mem.alloc(d_dead, (size_t)std::max<int64_t>(var_base[N], 1)));
        HIP_OR(hipMemcpyAsync(d_dead, dead, (size_t)var_base[N], hipMemcpyHostToDevice, st));
    }
    HIP_OR(hipMemcpyAsync(d_rm, rm.data(), sizeof(RefMeta) * N, hipMemcpyHostToDevice, st));
    HIP_OR(hipMemcpyAsync(d_vb, var_base.data(), sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice, st));
    RefArgs a;
    a.meta = d.meta;
    a.lits = d.lits;
    a.coff = d.coff;
    a.rmeta = d_rm;
    a.dead = d_dead;
    a.var_base = d_vb;
    a.ws = d_ws;
    a.n_inst = N;
    a.max_iterations = max_iterations;
    a.status = d_status;
    a.model_rows = d_models;
    a.stats = d_stats;
    HIP_OR(launch_solve_ref(a, st));
    HIP_OR(hipMemcpyAsync(status, d_status, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st));
    if (model_rows)
        HIP_OR(hipMemcpyAsync(model_rows, d_models, sizeof(uint32_t) * d.total_words, hipMemcpyDeviceToHost, st));
    if (stats) HIP_OR(hipMemcpyAsync(stats, d_stats, sizeof(int64_t) * N * 5, hipMemcpyDeviceToHost, st));
    HIP_OR(hipStreamSynchronize(st));
    return GPSAT_OK;
}

// solve_core.hpp -- the body of gpsat_solve, templated on where its device memory comes
// from (DevMem: hipMalloc per call, freed on return; a provider with kNoAlloc = true hands
// out pieces of a fixed region instead and refuses the parts of a solve that allocate).
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "../../include/gpsat.h"
#include "gpsat_internal.hpp"
#include "solver.hpp"

namespace gpsat {
// Per-device cache of device blocks that earlier calls no longer use, by rounded size
// (at most 1 GiB per device is kept; the rest is freed).  A solve repeated on the same shape
// then runs without hipMalloc/hipFree, which cost ~1 ms of a 12 ms solve of 20,000 instances.
void* devmem_cache_take(int dev, size_t bytes);
bool devmem_cache_give(int dev, void* p, size_t bytes);   // false: not kept (the caller frees)
size_t devmem_round(size_t bytes);

// Device memory of one call.  Pieces come from the cache or hipMalloc; when the call returns
// they go back to the cache if the call marked its device work finished (mark_synced), and
// are freed otherwise (hipFree waits for work still using them).
struct DevMem {
    struct Piece {
        void* p;
        size_t bytes;
    };
    std::vector<Piece> ptrs;
    int dev = -1;
    bool synced = false;
    ~DevMem()
    {
        for (const Piece& q : ptrs)
            if (!synced || !devmem_cache_give(dev, q.p, q.bytes)) (void)hipFree(q.p);
    }
    template <class T>
    hipError_t alloc(T*& p, size_t count)
    {
        if (dev < 0) {
            hipError_t e = hipGetDevice(&dev);
            if (e != hipSuccess) return e;
        }
        const size_t bytes = devmem_round(std::max<size_t>(count, 1) * sizeof(T));
        void* q = devmem_cache_take(dev, bytes);
        if (!q) {
            hipError_t e = hipMalloc(&q, bytes);
            if (e != hipSuccess) {
                p = nullptr;
                return e;
            }
        }
        ptrs.push_back({q, bytes});
        p = (T*)q;
        return hipSuccess;
    }
    void mark_synced() { synced = true; }
    static constexpr bool kNoAlloc = false;
    size_t remaining() const { return SIZE_MAX; }
};

inline int hip_fail(const char* what, hipError_t e)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? GPSAT_ERR_NOMEM : GPSAT_ERR_HIP;
}

#define HIP_OR(expr)                                  \
    do {                                              \
        hipError_t _e = (expr);                       \
        if (_e != hipSuccess) return hip_fail(#expr, _e); \
    } while (0)

// The solve itself, on the device memory Mem provides (hipMalloc, or the static scratch).
template <class Mem>
inline int solve_core(gpsat_batch* b, const gpsat_solve_opts& o, const int32_t* job_inst, const int32_t* cube_off,
               const int32_t* cube_lits, int32_t n_jobs, const float* act_init, const int8_t* phase_init,
               int32_t* status, uint32_t* model_rows, int32_t* verified, int64_t* job_stats, int32_t* var_bumps,
               gpsat_solve_info* info, Mem& mem, hipEvent_t ev[3])
{
    DeviceBatch& d = b->d;
    const int32_t N = d.n_inst;
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

    SolveArgs a{};
    SolvePlan plan{};
    if (int rc = solve_plan(d, o, n_jobs, a, plan)) return rc;
    const bool epochs = o.epoch_conflicts > 0;
    if (Mem::kNoAlloc && (epochs || (o.learned_out && o.learned_cap > 0))) {
        set_error("gpsat_solve_noalloc: restart epochs and the learned-clause dump need allocations");
        return GPSAT_ERR_ARG;
    }

    // ---- device buffers
    JobDesc* d_jobs;
    int32_t *d_cubes, *d_head, *d_status, *d_refuted, *d_undef, *d_ncubes, *d_ws = nullptr, *d_bumps = nullptr;
    int64_t *d_vbase, *d_jstats;
    uint32_t* d_models;
    float* d_act = nullptr;
    int8_t* d_phase = nullptr;
    int32_t* d_counts = nullptr;
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
    uint32_t* d_gtab = nullptr;   // mask tables of the model check for instances too large for LDS
    if (o.verify) {
        HIP_OR(mem.alloc(d_counts, (size_t)N * 3));
        if (Mem::kNoAlloc && eval_bcp_gtab_words(d) > 0) HIP_OR(mem.alloc(d_gtab, eval_bcp_gtab_words(d)));
    }
    if (var_bumps) HIP_OR(mem.alloc(d_bumps, (size_t)var_base[N]));
    if (act_init) HIP_OR(mem.alloc(d_act, (size_t)var_base[N]));
    if (phase_init) HIP_OR(mem.alloc(d_phase, (size_t)var_base[N]));
    if (plan.mode != 1) {
        // the workspace slices come last: a fixed region holds as many as are left room for
        const size_t slice = 4 * (size_t)a.ws_ints_per_wave;
        if ((size_t)plan.waves * slice > mem.remaining()) {
            const int fit = (int)(mem.remaining() / slice) / plan.wpb * plan.wpb;
            if (fit < plan.wpb) {
                set_error("gpsat_solve_noalloc: the solver workspace does not fit the scratch region");
                return GPSAT_ERR_CAPACITY;
            }
            set_waves(plan, fit);
        }
        HIP_OR(mem.alloc(d_ws, (size_t)plan.waves * a.ws_ints_per_wave));
    }
    a.chunk = plan.chunk;
    const int waves = plan.waves;
    if (act_init) HIP_OR(hipMemcpyAsync(d_act, act_init, sizeof(float) * var_base[N], hipMemcpyHostToDevice, st));
    if (phase_init) HIP_OR(hipMemcpyAsync(d_phase, phase_init, var_base[N], hipMemcpyHostToDevice, st));
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

    int32_t epochs_run = 1;
    int32_t* h_remaining = nullptr;
    if (epochs) {
        if (!(o.epoch_growth >= 1.0f) || o.max_epochs < 1 || o.max_epochs > 64) {
            set_error("invalid epoch options");
            return GPSAT_ERR_ARG;
        }
        int32_t *d_active, *d_nact, *d_epoch;
        int64_t* d_budget;
        float* d_astore;
        int8_t* d_pstore;
        HIP_OR(mem.alloc(d_active, 2 * (size_t)n_jobs));
        HIP_OR(mem.alloc(d_nact, 2));
        HIP_OR(mem.alloc(d_epoch, 1));
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

    HIP_OR(hipEventRecord(ev[0], st));
    if (!epochs) {
        HIP_OR(launch_solve(a, plan.mode, waves, plan.lds_bytes, st));
    } else {
        // one restart epoch = one replay of this graph: empty the queue, run the persistent
        // solve over the open jobs, re-queue the stopped ones, publish how many remain
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int32_t* d_remaining_ptr = nullptr;
        hipError_t ae = mem.alloc(d_remaining_ptr, 1);
        if (ae != hipSuccess) {
            (void)hipHostFree(h_remaining);
            return hip_fail("epoch counter", ae);
        }
        HIP_OR(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t ce = launch_epoch_begin(a, st);
        if (ce == hipSuccess) ce = launch_solve(a, plan.mode, waves, plan.lds_bytes, st);
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
    HIP_OR(hipEventRecord(ev[1], st));
    // re-verify every model: clause counts of the model rows (one row per instance)
    if (o.verify)
        HIP_OR(launch_eval_bcp(d, d_models, 1, d_counts, nullptr, nullptr, nullptr, nullptr, true, false, st, d_gtab));
    HIP_OR(hipEventRecord(ev[2], st));
    HIP_OR(hipStreamSynchronize(st));
    mem.mark_synced();   // every later device access below is a synchronous copy
    float ms_solve = 0, ms_verify = 0;
    (void)hipEventElapsedTime(&ms_solve, ev[0], ev[1]);
    (void)hipEventElapsedTime(&ms_verify, ev[1], ev[2]);

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
            t = (int64_t)((double)t * 1e6 / (double)plan.clock_khz);
        }
    if (job_stats) memcpy(job_stats, js.data(), sizeof(int64_t) * (size_t)n_jobs * JS_COUNT);
    if (var_bumps) HIP_OR(hipMemcpy(var_bumps, d_bumps, sizeof(int32_t) * var_base[N], hipMemcpyDeviceToHost));

    gpsat_solve_info inf{};
    inf.kernel_ms = ms_solve;
    inf.verify_ms = ms_verify;
    inf.waves = waves;
    inf.clause_mem = plan.mode;
    inf.lds_bytes = (int64_t)plan.lds_bytes;
    inf.workspace_bytes = plan.mode == 1 ? 0 : (int64_t)waves * a.ws_ints_per_wave * 4;
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

}  // namespace gpsat

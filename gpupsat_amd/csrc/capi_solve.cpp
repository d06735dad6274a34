// capi_solve.cpp -- gpsat_solve (launch plan, job table set-up, persistent solve launch,
// GPU re-verification of every SAT model with the clause-evaluation kernel) and the
// trajectory-mode solve gpsat_solve_ref.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpsat.h"
#include "gpsat_internal.hpp"
#include "solve_core.hpp"
#include "solver.hpp"

using namespace gpsat;

namespace gpsat {

namespace {
std::mutex g_cache_mu;
std::map<int, std::map<size_t, std::vector<void*>>> g_cache;   // device -> size -> free blocks
std::map<int, size_t> g_cache_bytes;
constexpr size_t kCacheMaxBytes = (size_t)1 << 30;
std::map<int, hipDeviceProp_t> g_props;
}  // namespace

size_t devmem_round(size_t bytes)
{
    if (bytes <= ((size_t)1 << 20)) return (bytes + 4095) & ~(size_t)4095;
    size_t p2 = (size_t)1 << 20;
    while (p2 < bytes) p2 <<= 1;
    const size_t step = p2 / 8;
    return (bytes + step - 1) / step * step;
}

void* devmem_cache_take(int dev, size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto& m = g_cache[dev];
    auto it = m.find(bytes);
    if (it == m.end() || it->second.empty()) return nullptr;
    void* p = it->second.back();
    it->second.pop_back();
    g_cache_bytes[dev] -= bytes;
    return p;
}

bool devmem_cache_give(int dev, void* p, size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    if (g_cache_bytes[dev] + bytes > kCacheMaxBytes) return false;
    g_cache[dev][bytes].push_back(p);
    g_cache_bytes[dev] += bytes;
    return true;
}

// Device properties, queried once per device.
static hipError_t device_props(int dev, hipDeviceProp_t& out)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_props.find(dev);
    if (it == g_props.end()) {
        hipDeviceProp_t prop;
        hipError_t e = hipGetDeviceProperties(&prop, dev);
        if (e != hipSuccess) return e;
        it = g_props.emplace(dev, prop).first;
    }
    out = it->second;
    return hipSuccess;
}

// The launch plan of one solve of n_jobs jobs under options o: the per-wave layout and
// solver options (into a) and the shape (where the job state sits, persistent wavefronts,
// jobs per queue grab).  gpsat_solve and the autotuner both use it, so a shape the
// autotuner measures is the shape a full run with the same options launches.
int solve_plan(const DeviceBatch& d, const gpsat_solve_opts& o, int32_t n_jobs, SolveArgs& a, SolvePlan& p)
{
    if (o.learnt_cap < 2 || o.restart_first < 1 || !(o.restart_factor >= 1.0f) || !(o.var_decay > 0.0f) ||
        o.var_decay > 1.0f || o.clause_mem < 0 || o.clause_mem > 4) {
        set_error("invalid solver options");
        return GPSAT_ERR_ARG;
    }
    int32_t max_L = 0;
    for (int32_t i = 0; i < d.n_inst; i++) max_L = std::max(max_L, d.h_coff[d.h_meta[i].coff_base + d.h_meta[i].m]);
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
    p.clock_khz = clock_khz;
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
    const size_t state_b = solve_state_lds_bytes(d.max_n);
    const size_t lds_all = state_b + 4 * (size_t)((clause_ints + 63) & ~(int64_t)63);

    hipDeviceProp_t prop;
    hipError_t e = device_props(d.device, prop);
    if (e != hipSuccess) return hip_fail("hipGetDeviceProperties", e);
    p.cus = prop.multiProcessorCount;
    const size_t lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 65536;
    const size_t lds_block_max = prop.sharedMemPerBlock ? prop.sharedMemPerBlock : 65536;
    const size_t lds_hyb = state_b + 4 * (size_t)(a.wo_cap + a.wpair_cap - 2 * o.learnt_cap);
    int mode;
    if (state_b > lds_block_max) mode = 4;   // per-var state alone exceeds the LDS: all in HBM
    else if (o.clause_mem == 1) mode = lds_all <= lds_block_max ? 1 : 2;
    else if (o.clause_mem == 2) mode = 2;
    else if (o.clause_mem == 3) mode = lds_hyb <= lds_block_max ? 3 : 2;
    else if (o.clause_mem == 4) mode = 4;
    else {
        // automatic: with few jobs the grid cannot fill the GPU and each job's latency is
        // what counts, so the clause part goes to LDS as far as it fits; with many, LDS only
        // while it leaves room for enough resident wavefronts
        const bool few_jobs = n_jobs <= 4 * prop.multiProcessorCount;
        if (lds_all <= (few_jobs ? lds_block_max : std::min<size_t>(48 * 1024, lds_block_max))) mode = 1;
        else if (few_jobs && lds_hyb <= lds_block_max) mode = 3;
        else mode = 2;
    }
    const int64_t slice_ints = clause_ints + (mode == 4 ? (int64_t)state_b / 4 : 0);
    a.ws_ints_per_wave = (slice_ints + 63) & ~(int64_t)63;
    p.mode = mode;
    p.lds_bytes = mode == 1 ? lds_all : mode == 3 ? lds_hyb : mode == 4 ? 0 : state_b;
    a.lds_per_wave = (int64_t)p.lds_bytes;
    int waves = o.waves;
    if (waves <= 0) {
        const size_t per_cu_lds = p.lds_bytes ? lds_cu / p.lds_bytes : (size_t)4 * kSolveWavesPerEu;
        const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4 * kSolveWavesPerEu, per_cu_lds));
        waves = prop.multiProcessorCount * per_cu;
    }
    p.wpb = solve_waves_per_block(p.lds_bytes);
    p.req_chunk = o.chunk;
    p.n_jobs = n_jobs;
    set_waves(p, waves);
    return GPSAT_OK;
}

// Persistent wavefronts (at most one per job, whole workgroups: surplus wavefronts find
// the queue empty and exit) and the chunk that goes with them.
void set_waves(SolvePlan& p, int waves)
{
    waves = std::max(1, std::min(waves, std::max(p.n_jobs, 1)));
    p.waves = (waves + p.wpb - 1) / p.wpb * p.wpb;
    // chunks of consecutive jobs let cubes of one instance reuse its watch lists; a few
    // chunks per wavefront keep the queue balanced
    p.chunk = (int32_t)std::max<int64_t>(1, std::min<int64_t>(64, (int64_t)p.n_jobs / ((int64_t)p.waves * 4)));
    if (p.req_chunk > 0) p.chunk = p.req_chunk;
}

}  // namespace gpsat

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
    if (b->d.n_inst == 0) return GPSAT_OK;
    DevMem mem;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int rc = GPSAT_OK;
    for (int k = 0; k < 3 && rc == GPSAT_OK; k++) {
        hipError_t e = hipEventCreate(&ev[k]);
        if (e != hipSuccess) rc = hip_fail("hipEventCreate", e);
    }
    if (rc == GPSAT_OK)
        rc = solve_core(b, o, job_inst, cube_off, cube_lits, n_jobs, act_init, phase_init, status, model_rows, verified,
                        job_stats, var_bumps, info, mem, ev);
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
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
    DeviceBatch& d = b->d;
    const int32_t N = d.n_inst;
    if (N == 0) return GPSAT_OK;
    std::vector<RefMeta> rm(N);
    std::vector<int64_t> var_base(N + 1, 0);
    int64_t ws = 0;
    for (int32_t i = 0; i < N; i++) {
        const InstMeta& mi = d.h_meta[i];
        const int32_t* co = d.h_coff.data() + mi.coff_base;
        int32_t largest = 0;
        for (int32_t c = 0; c < mi.m; c++) largest = std::max(largest, co[c + 1] - co[c]);
        // max_implication_per_var = max(largest clause, MIN_IMPLICATION_PER_VAR = 100)
        // (src/SATSolver/main.cu:148, Configs.cuh:55); the graph holds that many + 1 edges per vertex
        rm[i].ws = ws;
        rm[i].nw = mi.m + 21;
        rm[i].ecap = std::max(largest, 100) + 1;
        ws += (solve_ref_ws_ints(mi.n, mi.m, rm[i].ecap) + 15) & ~(int64_t)15;
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
    mem.mark_synced();
    return GPSAT_OK;
}

// solver.hpp -- structures shared by the device CDCL solver and its host launcher.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace gpsat {

// Solver options (the strategy surfaces of the reference's compile-time Configs.cuh,
// chosen at run time here).
// k_solve launch shape: independent wavefronts per workgroup, and the wavefronts per SIMD
// the register allocation is held to (8 -> at most 64 VGPRs per lane)
constexpr int kSolveWavesPerBlock = 4;
constexpr int kSolveWavesPerEu = 8;

struct SolveOptsDev {
    int32_t decision_mode;    // 0: VSIDS + phase saving; 1: reference scheme (last free var, positive)
    int32_t restart_first;    // conflicts before the first restart (GEOMETRIC_CONFLICTS_BEFORE_RESTART)
    float restart_factor;     // geometric growth (GEOMETRIC_RESTART_INCREASE_FACTOR)
    int32_t learnt_cap;       // learned clauses held per job
    int32_t larena_cap;       // int32 slots of learned-literal storage per job
    int32_t watch_slack;      // learned-clause watch slots per literal
    int64_t max_conflicts;    // per-job conflict budget (<= 0: none) -> UNDEF
    int64_t max_decisions;    // per-job decision cap (<= 0: none) -> UNDEF (MAX_ITERATIONS analogue)
    float var_decay;          // VSIDS activity decay per conflict
    int32_t phase_init;       // initial phase when no hint is given: 0 false, 1 true
    int64_t max_ticks;        // per-job wall-clock budget in device clock ticks (<= 0: none) -> UNDEF
    int64_t deadline_ticks;   // whole-solve budget in ticks after the first wavefront starts (<= 0: none)
    int64_t inst_conflicts;   // conflict pool of each instance over all its jobs (<= 0: none)
    int32_t telemetry;        // 1: accumulate per-phase ticks
};

struct JobDesc {
    int32_t inst;
    int32_t cube_off;
    int32_t cube_len;
    int32_t order;            // position of the job in the caller's job list
};

enum JobStat {
    JS_STATUS = 0, JS_CONFLICTS, JS_DECISIONS, JS_PROPS, JS_LEARNED, JS_RESTARTS, JS_LEVELMAX, JS_TICKS,
    // per-phase ticks (telemetry): set-up, BCP, analysis + learning, backtrack, decide
    JS_PH_SETUP, JS_PH_BCP, JS_PH_ANALYZE, JS_PH_BACKTRACK, JS_PH_DECIDE, JS_COUNT
};
enum Phase { PH_SETUP = 0, PH_BCP, PH_ANALYZE, PH_BACKTRACK, PH_DECIDE, PH_COUNT };

struct SolveArgs {
    const InstMeta* meta;
    const int32_t* lits;
    const int32_t* coff;
    const JobDesc* jobs;
    const int32_t* cube_lits;
    int32_t n_jobs;
    int32_t n_inst;
    const int64_t* var_base;        // [n_inst] first var slot of each instance in per-var arrays
    const float* act_init;          // [sum n] initial VSIDS scores (null: zero)
    const int8_t* phase_init;       // [sum n] initial phases 0/1 (null: opts.phase_init)
    int32_t* queue_head;            // device work-queue counter (counts chunks)
    unsigned long long* t_start;    // wall-clock tick at which the first wavefront started (0: not yet)
    int32_t chunk;                  // consecutive jobs handed out per queue grab
    int32_t* inst_status;           // [n_inst] 2 undecided, 0 SAT, 1 UNSAT
    int32_t* inst_refuted;          // [n_inst] refuted cubes
    int32_t* inst_undef;            // [n_inst] cubes stopped by a budget
    unsigned long long* inst_conf;  // [n_inst] conflicts spent by the instance's jobs so far
    const int32_t* inst_ncubes;     // [n_inst]
    uint32_t* model_rows;           // [total_words] model of each SAT instance (literal bitmap)
    int64_t* job_stats;             // [n_jobs * JS_COUNT]
    int32_t* var_bumps;             // [sum n] conflict-analysis participation per var (null: off)
    // restart epochs: the solve is a graph replayed once per epoch (all null otherwise)
    const int32_t* epoch;           // device epoch counter
    const int32_t* active;          // [2 * n_jobs] positions of this epoch's jobs, by parity
    const int32_t* n_active;        // [2] their numbers
    const int64_t* epoch_budget;    // conflict budget of every epoch
    float* act_store;               // [sum n] activities and phases kept by stopped jobs
    int8_t* phase_store;
    int32_t* ws;                    // global workspace
    int64_t ws_ints_per_wave;
    int64_t lds_per_wave;           // bytes of each wavefront's LDS region
    // learned-clause dump (testing aid; all null otherwise)
    int32_t* learned_out;
    int64_t learned_cap;
    unsigned long long* learned_pos;    // slots reserved by blocks that fit
    unsigned long long* learned_need;   // slots all blocks asked for
    // per-wave clause-part layout (ints), identical in LDS and global mode
    int32_t wo_cap, wl_cap, wpair_cap, lmeta_cap;
    int32_t max_n;
    int32_t all3;                   // every original clause has 3 literals (clause c at lits + 3c)
    SolveOptsDev o;
};

// Trajectory mode (solve_ref.hip): one instance per lane, the reference's own search.
struct RefMeta {
    int64_t ws;          // first int of the instance's workspace slice
    int32_t nw;          // watch-record slots: clauses + the 20 learned + 1
    int32_t ecap;        // backward edges per graph vertex (max implication per var + 1)
};

struct RefArgs {
    const InstMeta* meta;
    const int32_t* lits;
    const int32_t* coff;
    const RefMeta* rmeta;
    const uint8_t* dead;         // per var, instances back to back (null: none)
    const int64_t* var_base;     // [n_inst]
    int32_t* ws;
    int32_t n_inst;
    int32_t max_iterations;      // <= 0: none
    int32_t* status;             // [n_inst]
    uint32_t* model_rows;        // [total_words]
    int64_t* stats;              // [n_inst * 5]: decisions, conflicts, learned, iterations, restarts
};

int64_t solve_ref_ws_ints(int32_t n, int32_t m, int32_t ecap);
hipError_t launch_solve_ref(const RefArgs& a, hipStream_t st);

size_t solve_state_lds_bytes(int32_t max_n);

// Launch plan of one solve (solve_plan in capi_solve.cpp): mode 1 all job state in LDS,
// 2 clause part in HBM, 3 original clauses' watch structures in LDS and learned clauses in
// HBM, 4 all in HBM; the persistent wavefronts, jobs per queue grab, and LDS per wavefront.
struct SolvePlan {
    int mode;
    int waves;
    int chunk;
    int wpb;             // wavefronts per workgroup
    int req_chunk;       // the caller's chunk (0: automatic)
    int32_t n_jobs;
    int cus;
    int clock_khz;
    size_t lds_bytes;
};
hipError_t launch_epoch_begin(const SolveArgs& a, hipStream_t st);
hipError_t launch_epoch_end(const SolveArgs& a, int32_t* remaining, hipStream_t st);
hipError_t launch_solve(const SolveArgs& a, int clause_mem, int waves, size_t lds_bytes, hipStream_t st);
int solve_waves_per_block(size_t lds_bytes);

}  // namespace gpsat

struct gpsat_solve_opts;

namespace gpsat {
struct DeviceBatch;
int solve_plan(const DeviceBatch& d, const gpsat_solve_opts& o, int32_t n_jobs, SolveArgs& a, SolvePlan& p);
void set_waves(SolvePlan& p, int waves);

}  // namespace gpsat

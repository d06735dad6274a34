/*
 * gpsat.h -- C ABI of the MI355X SAT engine (libgpsat.so).
 *
 * Plain pointers and sizes only; every entry point returns an int status
 * (GPSAT_OK = 0, negative = error; gpsat_last_error() gives the message) and never
 * exits the process.  A batch handle owns one HIP stream; calls on different
 * handles may run from different host threads, calls on one handle may not.
 *
 * Literal code l = 2*v + s with s = 1 for the positive literal
 * (reference src/SATSolver/SolverTypes.cu:6-11).  Verdict codes follow the
 * reference's sat_status order SAT = 0, UNSAT = 1, UNDEF = 2
 * (src/SATSolver/SolverTypes.cuh:125).
 *
 * Assignment rows: one row per (instance, assignment) is a bitmap indexed by
 * literal code, W_i = ceil(2 n_i / 32) uint32 words; bit l set <=> literal l is
 * true (var v: 00 unassigned, 01 false, 10 true, 11 contradictory).  With A rows
 * per instance, the rows of instance i start at word A * sum_{j<i} W_j
 * (gpsat_batch_row_words gives the total).
 *
 * Suffix _dev: pointers are device memory (HBM), the call is asynchronous on the
 * batch stream.  Without the suffix: host memory, the call is synchronous.
 */
#ifndef GPSAT_H
#define GPSAT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPSAT_OK 0
#define GPSAT_ERR_ARG (-1)
#define GPSAT_ERR_HIP (-2)
#define GPSAT_ERR_NOMEM (-3)
#define GPSAT_ERR_CAPACITY (-4)
#define GPSAT_ERR_NODEVICE (-5)
#define GPSAT_ERR_PARSE (-6)

#define GPSAT_SAT 0
#define GPSAT_UNSAT 1
#define GPSAT_UNDEF 2

/* int64 fields per job in gpsat_solve's job_stats */
#define GPSAT_JOB_STATS 13

typedef struct gpsat_batch gpsat_batch;

/* Last error message of the calling thread ("" if none). */
const char *gpsat_last_error(void);
/* Library version, "v0.0.1" semantics of the reference's --version (src/FileManager/Version.h:3). */
const char *gpsat_version(void);
/* Number of visible GPUs (0 when none). */
int gpsat_device_count(int32_t *count);
int gpsat_device_synchronize(void);

/* ---- device memory and timing helpers (plumbing for callers without a HIP binding) */
int gpsat_dev_alloc(int64_t bytes, void **ptr);
int gpsat_dev_free(void *ptr);
int gpsat_memcpy_h2d(void *dst_dev, const void *src_host, int64_t bytes);
int gpsat_memcpy_d2h(void *dst_host, const void *src_dev, int64_t bytes);
int gpsat_memset_dev(void *dst_dev, int32_t value, int64_t bytes);
int gpsat_event_create(void **ev);
int gpsat_event_destroy(void *ev);
int gpsat_event_record(void *ev, gpsat_batch *b);
int gpsat_event_elapsed_ms(void *start, void *stop, float *ms);
/* Streaming device copy of bytes (a multiple of 16) on the batch stream (or the null
 * stream): the practical-peak and counter-scaling probe; no reference counterpart. */
int gpsat_copy_dev(gpsat_batch *b, void *dst_dev, const void *src_dev, int64_t bytes);

/* ---- launch-shape tuning (process-wide; no reference counterpart).  Keys:
 *   "eval_fast3"        1: evaluate batches whose clauses all have 3 literals, n <= 256 and
 *                       m <= 1024 with the register-resident kernel (default); 0: general kernel
 *   "eval_stage_arena"  1: the general kernel copies each clause arena to LDS first (default 0)
 *   "eval_waves_per_block"  instances (wavefronts) per workgroup of the eval/BCP kernels:
 *                       1, 2 or 4 (default 4)
 *   "eval_fast3_small64"  3-literal kernel: instances with at most this many rows open after
 *                       the evaluation run their BCP passes on the 64-bit table instead of a
 *                       compacted 32-bit one (default 8; 0: always compact)
 *   "eval_pkk_waves"    wavefronts sharing one instance in the packed general kernel, 1-4 (default 4)
 *   "eval_pkk_compact"  packed general kernel: from the second BCP pass on, long instances with
 *                       at most this many rows still moving continue on 32-bit masks of those
 *                       rows, 0-32 (default 32; 0: never)
 *   "eval_pkk"          1: evaluate batches of clauses with at most 6 literals and n <= 496
 *                       (that the register-resident kernel does not take) with the packed
 *                       general kernel (default); 0: the CSR general kernel
 * The GPSAT_TUNE_EVAL_FAST3 / GPSAT_TUNE_EVAL_STAGE / GPSAT_TUNE_EVAL_WPB /
 * GPSAT_TUNE_EVAL_FAST3_SMALL64 / GPSAT_TUNE_EVAL_PKK / GPSAT_TUNE_EVAL_PKK_WAVES /
 * GPSAT_TUNE_EVAL_PKK_COMPACT environment variables set the initial values.  Returns GPSAT_ERR_ARG
 * for an unknown key. */
int gpsat_tune_set(const char *key, int32_t value);
int gpsat_tune_get(const char *key, int32_t *value);

/* ---- formula batches.
 * Replaces the reference's per-formula device arena CUDAClauseVec::alloc_and_copy_to_dev
 * (src/Utils/CUDAClauseVec.cu:85-118) with one CSR arena for many instances:
 * lits[clause_off[c] .. clause_off[c+1]) is clause c, instance i owns clauses
 * inst_clause_off[i] .. inst_clause_off[i+1], n_vars[i] variables.  Device memory: 4 B per
 * literal, 4 B per clause offset and 32 B per instance; when every clause has 3 literals,
 * n <= 256 and m <= 1024, also 4 B per clause of packed literal triples; otherwise, when
 * every clause has at most 6 literals and n <= 496, 8 B per clause of packed literal codes
 * and 1 B per instance. */
int gpsat_batch_create(const int32_t *lits, const int64_t *clause_off, const int64_t *inst_clause_off,
                       const int32_t *n_vars, int32_t n_inst, gpsat_batch **out);
int gpsat_batch_destroy(gpsat_batch *b);
int gpsat_batch_info(const gpsat_batch *b, int32_t *n_inst, int64_t *n_clauses, int64_t *n_lits, int32_t *max_vars);
int gpsat_batch_row_words(const gpsat_batch *b, int32_t A, int64_t *total_words);
int gpsat_batch_sync(gpsat_batch *b);
/* Which kernel an eval/BCP launch on this batch runs under the current tuning: 0 the
 * register-resident 3-literal kernel, 1 the packed general kernel, 2 the CSR general kernel
 * with its mask table in LDS, 3 the CSR general kernel with mask tables in HBM slices
 * (n too large for LDS).  No reference counterpart (diagnostics). */
int gpsat_batch_eval_kernel(const gpsat_batch *b, int32_t *kind);

/* ---- a4: clause evaluation, VariablesStateHandler::clause_status
 * (src/SATSolver/VariablesStateHandler.cu:180-206) for every clause under every
 * row.  counts[(i*A + a)*3 + {0,1,2}] = #SAT, #UNSAT, #UNIT.  1 <= A <= 64. */
int gpsat_clause_eval_dev(gpsat_batch *b, const uint32_t *rows_dev, int32_t A, int32_t *counts_dev);
int gpsat_clause_eval(gpsat_batch *b, const uint32_t *rows, int32_t A, int32_t *counts);

/* ---- a5: BCP fixpoint of each row, the closure the reference's watched-literal
 * BCP reaches when the row's literals are applied as a trail
 * (src/BCPStrategy/WatchedClausesList.cu:46-309, WatchedClausesList::new_decision).
 * conflict[i*A+a] = 1 on a conflict (contradictory row, an all-false clause, or
 * complementary implications); then out row = input row and props = 0.  Otherwise
 * out row = closure and props = number of implied literals.  1 <= A <= 64. */
int gpsat_bcp_fixpoint_dev(gpsat_batch *b, const uint32_t *rows_dev, int32_t A, uint32_t *out_rows_dev,
                           int32_t *conflict_dev, int32_t *props_dev);
int gpsat_bcp_fixpoint(gpsat_batch *b, const uint32_t *rows, int32_t A, uint32_t *out_rows, int32_t *conflict,
                       int32_t *props);
/* Both in one launch (clause counts of the input rows, then the fixpoint). */
int gpsat_eval_bcp_dev(gpsat_batch *b, const uint32_t *rows_dev, int32_t A, int32_t *counts_dev,
                       uint32_t *out_rows_dev, int32_t *conflict_dev, int32_t *props_dev);

/* ---- host-side input: DIMACS CNF, the reader of src/FileManager/CnfReader.cpp:30-135
 * (CnfManager::read_cnf).  n_vars is the highest var seen (CnfReader.cpp:117-123);
 * the header's counts are kept for the reference's mismatch warnings. */
typedef struct gpsat_formula gpsat_formula;
int gpsat_dimacs_read(const char *path, gpsat_formula **out);
int gpsat_dimacs_parse(const char *text, int64_t len, gpsat_formula **out);
int gpsat_formula_info(const gpsat_formula *f, int32_t *n_vars, int32_t *n_clauses, int64_t *n_lits,
                       int32_t *header_vars, int32_t *header_clauses);
/* lits[n_lits] and clause_off[n_clauses + 1] */
int gpsat_formula_arrays(const gpsat_formula *f, int32_t *lits, int64_t *clause_off);
int gpsat_formula_free(gpsat_formula *f);

/* ---- a3: preprocessing, FormulaData::copy_host_clauses_to_dev
 * (src/FileManager/FormulaData.cu:82-105): RepeatedLiteralsRemover then
 * UnaryClausesRemover.  Output buffers hold at least the input sizes; solved holds
 * up to m literals.  status: 0 SAT, 1 UNSAT (decided by preprocessing), 2 formula remains. */
int gpsat_preprocess(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n_vars, int32_t *out_lits,
                     int64_t *out_clause_off, int32_t *out_m, int32_t *solved, int32_t *n_solved, int32_t *status);

/* The same preprocessing for a whole batch on the device, one wavefront per instance
 * (no reference counterpart as a batch; semantics of FormulaData.cu:82-105 per instance).
 * Outputs (host buffers): status[n_inst] (0 SAT, 1 UNSAT, 2 UNDEF); solved_rows
 * (gpsat_batch_row_words(b, 1) words: the solved literals of each instance as a literal
 * bitmap, zero when it had no unit clause); the remaining formulas in gpsat_batch_create's
 * layout: out_lits[<= n_lits], out_clause_off[<= n_clauses + 1] (int64, global),
 * out_inst_clause_off[n_inst + 1]; an UNSAT instance keeps no clauses. */
int gpsat_preprocess_batch(gpsat_batch *b, int32_t *status, uint32_t *solved_rows, int32_t *out_lits,
                           int64_t *out_clause_off, int64_t *out_inst_clause_off);

/* ---- a14: cube generation, the JobChooser surface (src/JobsManager/JobChooser.cuh:21-31).
 * strategy 0: MaxClauseJobChooser distributing jobs per thread, 1: MaxClauseJobChooser
 * uniform (src/JobsManager/JobChooser.cu:52-133, VariableChooser.cu:23-40); 2:
 * SimpleJobChooser (src/JobsManager/SimpleJobChooser.cu:22-49).  out_vars holds up to
 * 15 vars; *depth gets their number. */
int gpsat_cube_vars(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n_vars, const int32_t *dead_vars,
                    int32_t n_dead, int64_t threads, int64_t blocks, int32_t strategy, int32_t *out_vars,
                    int32_t *depth);
/* The depth highest-scoring live vars (for example the learned network's scores);
 * no reference counterpart. */
int gpsat_cube_vars_scored(const float *scores, int32_t n_vars, const int32_t *dead_vars, int32_t n_dead,
                           int32_t depth, int32_t *out_vars, int32_t *out_depth);
/* All 2^depth cubes over vars, depth first with the positive literal first
 * (MaxClauseJobChooser::addJobs, JobChooser.cu:68-90): out_lits[2^depth * depth]. */
int gpsat_cubes(const int32_t *vars, int32_t depth, int32_t *out_lits);

/* ---- a6..a13, a15: the solve path.
 * Replaces the reference's solve kernels run_sequential / parallel_kernel
 * (src/SATSolver/Parallelizer.cu:180-278, launched from src/SATSolver/main.cu:211,256-269):
 * one wavefront per job, a persistent grid pulling jobs from a device queue.  A job is
 * an instance plus a cube of literals assumed true (JobsQueue::next_job, JobsQueue.cu:10-32);
 * an instance is SAT when one of its jobs is SAT and UNSAT only when every one of its
 * jobs is refuted.  The strategy surfaces of the reference (Configs.cuh) are options. */
typedef struct gpsat_solve_opts {
    int32_t decision;        /* 0: VSIDS + phase saving; 1: reference rule, last free var positive
                                (src/SATSolver/DecisionMaker.cu:45-55) */
    int32_t restart_first;   /* conflicts before the first restart (reference 100, Configs.cuh:104) */
    float restart_factor;    /* geometric growth (reference 1.3, Configs.cuh:105) */
    int32_t learnt_cap;      /* learned clauses held per job (reference 20, Configs.cuh:94) */
    int32_t larena_cap;      /* learned-literal slots per job (0: automatic) */
    int32_t watch_slack;     /* learned-clause watch slots per literal (0: automatic) */
    int64_t max_conflicts;   /* per-job conflict budget (<= 0: none), exceeded -> UNDEF */
    int64_t max_decisions;   /* per-job decision cap, MAX_ITERATIONS analogue (Configs.cuh:23) */
    float var_decay;         /* VSIDS decay per conflict */
    int32_t phase_init;      /* initial phase without a hint: 0 false, 1 true */
    double max_ms;           /* per-job wall-clock budget in ms (<= 0: none), exceeded -> UNDEF */
    int32_t waves;           /* persistent wavefronts, rounded up to whole workgroups of 4
                                (0: automatic, up to 32 per CU as LDS allows) */
    int32_t clause_mem;      /* clause part of the job state: 0 automatic, 1 LDS, 2 HBM, 3 the
                                original clauses' watch lists and pairs in LDS, the learned
                                clauses' storage in HBM.  Automatic: with at most 4 jobs per CU,
                                LDS, else split, else HBM, whichever fits first; with more, LDS
                                when it takes at most 48 KB per wavefront, else HBM */
    int32_t verify;          /* 1: re-check every SAT model with the clause-evaluation kernel */
    int32_t chunk;           /* jobs per queue grab (0: automatic); cubes of one instance in one
                                chunk reuse its watch lists on the wavefront */
    int64_t epoch_conflicts; /* > 0: restart epochs.  The solve becomes a HIP graph (queue reset,
                                persistent solve, re-queue of stopped jobs) replayed once per
                                epoch; in epoch e a job may use epoch_conflicts * growth^e
                                conflicts, and a stopped job keeps its activities and phases */
    float epoch_growth;      /* budget growth per epoch (default 2) */
    int32_t max_epochs;      /* epochs before the remaining jobs are UNDEF (default 16) */
    double deadline_ms;      /* whole-solve wall-clock budget in ms from the first wavefront's
                                start (<= 0: none): jobs running when it passes stop, jobs not
                                yet started are not run; both count as UNDEF */
    int64_t inst_conflicts;  /* per-instance conflict pool shared by all its jobs (<= 0: none):
                                once spent, its running jobs stop and the rest are skipped (UNDEF) */
    int32_t job_order;       /* queue order: 0 largest instances first (default: the grid ends on
                                short jobs), 1 smallest first (decides the most instances before a
                                deadline), 2 as given; cubes of one instance stay adjacent */
    int32_t telemetry;       /* 1: per-phase times in job_stats (set-up, BCP, analysis, backtrack,
                                decide), the phases of RuntimeStatistics (src/Statistics/
                                RuntimeStatistics.cu:246-360); 0: those fields stay 0 */
    int32_t *learned_out;    /* optional host buffer (testing aid; no reference counterpart): at
                                the end of each job that ran, its held learned clauses and its
                                level-0 literals (as unit clauses), as blocks [instance, slots,
                                then per clause its literal codes and -1]; each is implied by the
                                formula and the job's cube.  A block that does not fit is left out */
    int64_t learned_cap;     /* int32 slots of learned_out */
} gpsat_solve_opts;

typedef struct gpsat_solve_info {
    double kernel_ms;        /* solve kernel (HIP events on the batch stream) */
    double verify_ms;        /* model re-verification kernel */
    int32_t waves;
    int32_t clause_mem;      /* 1 LDS, 2 HBM, 3 split (see gpsat_solve_opts) */
    int64_t lds_bytes;       /* per wavefront */
    int64_t workspace_bytes;
    int32_t n_sat, n_unsat, n_undef, n_verified, n_errors;
    int32_t epochs;          /* graph replays in epoch mode (1 otherwise) */
    int32_t chunk;           /* jobs per queue grab that was used */
    int64_t learned_len;     /* slots written to opts.learned_out */
    int64_t learned_need;    /* slots all blocks needed (> learned_len: some were left out) */
} gpsat_solve_info;

int gpsat_solve_opts_default(gpsat_solve_opts *o);
/* job_inst: instance of each job (NULL: one job per instance, no cubes); cube_off[n_jobs+1]
 * and cube_lits: the literals of each job's cube.  act_init / phase_init: per-var
 * initial VSIDS scores and phases, instances back to back (NULL: none).  Outputs:
 * status[n_inst] (0 SAT, 1 UNSAT, 2 UNDEF); model_rows (gpsat_batch_row_words(b, 1)
 * words: one literal bitmap per instance, valid where SAT); verified[n_inst] (1 where a
 * SAT model was re-checked on the GPU); job_stats[n_jobs*GPSAT_JOB_STATS] (status, conflicts,
 * decisions, propagations, learned, restarts, max level, job time in ns, then with telemetry the
 * ns spent in set-up, BCP, conflict analysis + learning, backtracking, decisions); var_bumps (per var, how
 * often conflict analysis met it).  Every output may be NULL except status.
 * Device memory: the call's buffers come from hipMalloc or from blocks earlier solves left
 * behind; when a call ends they are kept for reuse (up to 1 GiB per device, held until the
 * process ends), so repeating a solve of one shape allocates nothing after the first call. */
int gpsat_solve(gpsat_batch *b, const gpsat_solve_opts *o, const int32_t *job_inst, const int32_t *cube_off,
                const int32_t *cube_lits, int32_t n_jobs, const float *act_init, const int8_t *phase_init,
                int32_t *status, uint32_t *model_rows, int32_t *verified, int64_t *job_stats, int32_t *var_bumps,
                gpsat_solve_info *info);

/* Allocation-free solve (no counterpart in the reference): gpsat_solve for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same arguments, outputs and verdicts as gpsat_solve; every
 * device buffer (job table, queue and instance counters, job statistics, model rows, the
 * solver workspace, the model check's counts) is carved from one static device region of
 * GPSAT_NOALLOC_SCRATCH_MIB MiB per GPU, which the library holds until the process ends; no
 * hipMalloc, hipFree or pinned allocation is made.  When the clause part of the job state
 * sits in HBM, fewer persistent wavefronts are launched if their slices would not fit.  A call
 * whose buffers exceed the region returns GPSAT_ERR_CAPACITY before anything runs; restart
 * epochs (epoch_conflicts > 0) and the learned-clause dump return GPSAT_ERR_ARG.  Every call
 * on a GPU shares that GPU's region: calls that overlap on one GPU race. */
#define GPSAT_NOALLOC_SCRATCH_MIB 64
int gpsat_solve_noalloc(gpsat_batch *b, const gpsat_solve_opts *o, const int32_t *job_inst, const int32_t *cube_off,
                        const int32_t *cube_lits, int32_t n_jobs, const float *act_init, const int8_t *phase_init,
                        int32_t *status, uint32_t *model_rows, int32_t *verified, int64_t *job_stats,
                        int32_t *var_bumps, gpsat_solve_info *info);

/* Trajectory mode: the reference's own sequential search per instance, step for step
 * (SATSolver::solve, src/SATSolver/SATSolver.cu:67-218, with the default strategies of
 * Configs.cuh), one instance per wavefront, over the reference's structures: per-variable watch
 * lists with the fewest-watched-clauses choice and the blocked list
 * (BCPStrategy/ClauseListStructure.cu:39-243), depth-first propagation of pending implications
 * (WatchedClausesList.cu:46-309), the implication graph and decision-scheme learning
 * (ConflictAnalysis/GraphAnalyzer.cu:31-138, GraphStructure.cu:79-330), Backtracker's backjump
 * (Backtracker.cu:37-75), the 20-clause round-robin repository, geometric restarts 100 x 1.3
 * and MAX_ITERATIONS (max_iterations > 0).  Clauses of fewer than two literals are not watched,
 * as in the reference's solver (its input has been preprocessed).  dead: per var (instances
 * back to back, may be NULL) variables the preprocessing solved, kept out of the free list as in
 * FormulaData/Results.  learned_cap and learned_lits are not used (the search holds at most the
 * reference's 20 learned clauses and the one being analysed); pass 0.  Outputs: status[n_inst]
 * (0 SAT, 1 UNSAT, 2 UNDEF, 4 a graph vertex needed more backward edges than the reference's
 * max_implication_per_var + 1, where the reference writes past its edge array), model_rows
 * (literal bitmaps, gpsat_batch_row_words(b, 1) words), stats[n_inst*5] (decisions, conflicts,
 * learned clauses, iterations, restarts; any output but status may be NULL). */
int gpsat_solve_ref(gpsat_batch *b, const uint8_t *dead, int32_t max_iterations, int32_t learned_cap,
                    int32_t learned_lits, int32_t *status, uint32_t *model_rows, int64_t *stats);

/* ---- launch-shape autotuner (no reference counterpart: the reference takes -b/-t and
 * fixes its strategies at compile time, src/SATSolver/Configs.cuh).
 * gpsat_autotune_solve: short trial solves of a sample of the jobs (whole instances, at
 * most sample_jobs jobs; 0: 4096) under each launch shape -- clause part in LDS, HBM or split,
 * 8/16/24/32 persistent wavefronts per CU, then 1/8/32/64 jobs per queue grab when there
 * are cubes -- each with deadline trial_ms (0: 300).  *best = *base with waves,
 * clause_mem and chunk of the shape with the most propagations per ms of kernel time.
 * Verdicts do not depend on the shape.  report (optional) receives one line per trial.
 * gpsat_autotune_eval: times reps launches (0: 10) of clause evaluation and/or BCP on the
 * caller's rows under each shape of every kernel the batch can take -- the 3-literal kernel
 * ("eval_waves_per_block" x "eval_fast3_small64"), the packed general kernel
 * ("eval_pkk_waves" x "eval_pkk_compact" 0/32) and the CSR kernel ("eval_stage_arena" x
 * "eval_waves_per_block") -- and leaves the tuning keys at the fastest.  Outputs do not
 * depend on the shape. */
int gpsat_autotune_solve(gpsat_batch *b, const gpsat_solve_opts *base, const int32_t *job_inst, const int32_t *cube_off,
                         const int32_t *cube_lits, int32_t n_jobs, int32_t sample_jobs, double trial_ms,
                         gpsat_solve_opts *best, char *report, int32_t report_len);
int gpsat_autotune_eval(gpsat_batch *b, const uint32_t *rows_dev, int32_t A, int32_t do_eval, int32_t do_bcp,
                        int32_t reps, char *report, int32_t report_len);

/* ---- a17: the learned branching network (no reference counterpart).  Message passing
 * over the literal-clause incidence graph of a batch: 64 hidden units, `rounds` rounds,
 * BF16 weights with FP32 accumulation on MFMA; per var y = (branching score, phase
 * logit).  A model holds FP32 master parameters (gpsat_gnn_param_count floats) and
 * Adam state; a graph holds one batch's incidence structure, activations and labels.
 * Labels: target score, phase (0/1) and phase mask per var, instances back to back. */
typedef struct gpsat_gnn_model gpsat_gnn_model;
typedef struct gpsat_gnn_graph gpsat_gnn_graph;
int gpsat_gnn_param_count(int32_t *n);
int gpsat_gnn_model_create(const float *theta, gpsat_gnn_model **out);
int gpsat_gnn_model_destroy(gpsat_gnn_model *m);
int gpsat_gnn_model_params(gpsat_gnn_model *m, float *theta);
int gpsat_gnn_graph_create(gpsat_batch *b, int32_t rounds, gpsat_gnn_graph **out);
int gpsat_gnn_graph_destroy(gpsat_gnn_graph *g);
int gpsat_gnn_graph_labels(gpsat_gnn_graph *g, const float *target, const float *phase, const float *mask);
/* y[V*2] */
int gpsat_gnn_forward(gpsat_gnn_model *m, gpsat_gnn_graph *g, float *y);
/* mean loss over vars and its gradient (no update); loss, grad and y may be NULL */
int gpsat_gnn_loss_grad(gpsat_gnn_model *m, gpsat_gnn_graph *g, float alpha, float *loss, float *grad, float *y);
/* one Adam step (beta 0.9 / 0.999) on the batch */
int gpsat_gnn_train_step(gpsat_gnn_model *m, gpsat_gnn_graph *g, float alpha, float lr, float *loss);

#ifdef __cplusplus
}
#endif
#endif

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

/* ---- formula batches.
 * Replaces the reference's per-formula device arena CUDAClauseVec::alloc_and_copy_to_dev
 * (src/Utils/CUDAClauseVec.cu:85-118) with one CSR arena for many instances:
 * lits[clause_off[c] .. clause_off[c+1]) is clause c, instance i owns clauses
 * inst_clause_off[i] .. inst_clause_off[i+1], n_vars[i] variables. */
int gpsat_batch_create(const int32_t *lits, const int64_t *clause_off, const int64_t *inst_clause_off,
                       const int32_t *n_vars, int32_t n_inst, gpsat_batch **out);
int gpsat_batch_destroy(gpsat_batch *b);
int gpsat_batch_info(const gpsat_batch *b, int32_t *n_inst, int64_t *n_clauses, int64_t *n_lits, int32_t *max_vars);
int gpsat_batch_row_words(const gpsat_batch *b, int32_t A, int64_t *total_words);
int gpsat_batch_sync(gpsat_batch *b);

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

#ifdef __cplusplus
}
#endif
#endif

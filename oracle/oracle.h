/*
 * oracle.h -- CPU restatement of the gpupsat solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gpupsat_amd/) links, loads
 * or calls this code.  It is used by tests/ as the checker, by
 * __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg as
 * the timed CPU restatement of the reference's sequential solver.
 *
 * Conventions (reference src/SATSolver/SolverTypes.cu:6-32):
 *   literal code l = 2*v + s, s = 1 means the positive literal; var(l) = l >> 1;
 *   negation is l ^ 1.  DIMACS x > 0 maps to 2*(x-1)+1 (src/FileManager/CnfReader.cpp:106-112).
 *   Verdict codes follow sat_status order SAT=0, UNSAT=1, UNDEF=2
 *   (src/SATSolver/SolverTypes.cuh:125).
 *
 * Assignment rows: one row per (instance, assignment) is a bitmap indexed by
 * literal code with W = ceil(2n/32) uint32 words: bit l set <=> literal l is
 * true.  For var v this is the 2-bit code at bits (2v, 2v+1):
 * 00 unassigned, 01 false, 10 true, 11 contradictory.
 *
 * Batch CSR (host form): lits[Ltot]; clause_off[Mtot+1] (int64, global literal
 * offsets); inst_clause_off[n_inst+1] (int64, global clause offsets); n_vars[n_inst].
 * Rows of instance i start at word A * sum_{j<i} W_j.
 */
#ifndef GPSAT_ORACLE_H
#define GPSAT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_SAT = 0, ORC_UNSAT = 1, ORC_UNDEF = 2 };

/* ---- a4: clause status, VariablesStateHandler::clause_status (VariablesStateHandler.cu:180-206) */
int orc_clause_status(const int32_t *lits, int32_t n, const uint32_t *row, int32_t *unit_lit);

/* Per (instance, assignment): counts[(i*A+a)*3 + {0,1,2}] = #SAT, #UNSAT, #UNIT clauses. */
void orc_eval_batch(int32_t n_inst, const int32_t *lits, const int64_t *clause_off,
                    const int64_t *inst_clause_off, const int32_t *n_vars,
                    const uint32_t *rows, int32_t A, int32_t *counts);

/* ---- a5: BCP fixpoint = unit-propagation closure of a trail.
 * Returns the conflict flag.  Non-conflict: out_row = closure, *props = |closure \ trail|.
 * Conflict (contradictory input, an all-false clause, or complementary implications):
 * out_row = in_row, *props = 0. */
int orc_up_closure(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                   const uint32_t *in_row, uint32_t *out_row, int32_t *props);

void orc_bcp_batch(int32_t n_inst, const int32_t *lits, const int64_t *clause_off,
                   const int64_t *inst_clause_off, const int32_t *n_vars,
                   const uint32_t *rows, int32_t A, uint32_t *out_rows,
                   int32_t *conflict, int32_t *props);

/* ---- a3: preprocessing, FormulaData::copy_host_clauses_to_dev (FormulaData.cu:82-105):
 * RepeatedLiteralsRemover (RepeatedLiteralsRemover.cu:25-62) then
 * UnaryClausesRemover (UnaryClausesRemover.cu:13-173).
 * Output buffers must hold at least the input sizes.  Returns the status after
 * preprocessing (ORC_UNDEF if the formula remains). */
int orc_preprocess(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                   int32_t *out_lits, int64_t *out_clause_off, int32_t *out_m,
                   int32_t *solved, int32_t *n_solved);

/* ---- a6..a13: the reference's sequential CDCL scheme (SATSolver.cu:67-218):
 * decision = last free var, positive (DecisionMaker.cu:45-55); propagation to a
 * fixpoint; decision-scheme ("last UIP" at the conflict level) learning
 * (GraphAnalyzer.cu:31-138) with backjump to the second-highest level
 * (ConflictAnalyzer.cu:105-151); 20-clause round-robin learned DB
 * (LearntClauseRepository.cu:15-33); geometric restarts 100 / x1.3
 * (GeometricRestartsManager.cu:3-31); optional iteration cap -> UNDEF
 * (Configs.cuh:23, SATSolver.cu:182-188).  dead[v] != 0 removes v from the
 * free list (VariablesStateHandler.cu:4-33).
 * model[v]: 1 true, 0 false, -1 unassigned.  stats: decisions, conflicts,
 * learned, iterations, restarts. */
int orc_solve_ref(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                  const uint8_t *dead, int32_t max_iterations, int8_t *model, int64_t *stats);

/* The reference's sequential path (preprocessing, then the solver with solved vars as
 * dead vars) for every instance of a batch; status[i] gets the verdict. */
void orc_solve_ref_batch(int32_t n_inst, const int32_t *lits, const int64_t *clause_off,
                         const int64_t *inst_clause_off, const int32_t *n_vars, int32_t max_iterations,
                         int32_t *status);

/* Independent checkers for verdicts (no reference counterpart). */
int orc_solve_dpll(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                   int8_t *model, int64_t max_nodes);
int orc_solve_brute(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                    int8_t *model);
/* Textbook CDCL (1-UIP, VSIDS, phase saving, Luby restarts); conflict budget <0 = none. */
int orc_solve_cdcl(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                   int8_t *model, int64_t max_conflicts, int64_t *stats);

/* 1 iff every clause has a true literal under model (values 1/0; -1 counts as false). */
int orc_verify_model(const int32_t *lits, const int64_t *clause_off, int32_t m, const int8_t *model);

/* ---- a14: MaxClauseJobChooser (JobChooser.cu:52-133) + VariableChooser (VariableChooser.cu:23-40).
 * strategy 0 = distribute jobs per thread, 1 = uniform.  Ties in the score sort
 * are broken by ascending var.  Returns the depth d; out_vars gets d vars. */
int orc_choose_cube_vars(const int32_t *lits, const int64_t *clause_off, int32_t m, int32_t n,
                         int32_t n_dead, int64_t threads, int64_t blocks, int32_t strategy,
                         int32_t *out_vars);
/* Cube c (0 <= c < 2^d) holds literal j positive iff bit (d-1-j) of c is 0
 * (JobChooser.cu:68-90 DFS, positive first).  out: 2^d * d literal codes. */
void orc_enumerate_cubes(const int32_t *vars, int32_t d, int32_t *out);

/* ---- a12: restart thresholds: t0 = first, t_{k+1} = (int)(t_k * (float)factor). */
void orc_restart_schedule(int32_t first, float factor, int32_t k, int32_t *out);

#ifdef __cplusplus
}
#endif
#endif

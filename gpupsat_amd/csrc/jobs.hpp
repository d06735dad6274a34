// jobs.hpp -- host-side job generation: the reference's JobChooser plugin surface
// (src/JobsManager/JobChooser.cuh:21-31) and JobsQueue (src/SATSolver/JobsQueue.cuh:24-42).
//
// A job is a cube: a list of literals assumed true.  JobsQueue collects them on the
// host; gpsat_solve ships the whole table to the device queue in one copy (the
// reference allocates and copies every cube separately, src/SATSolver/JobsQueue.cu:65-78).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace gpsat {

enum class ChoosingStrategy { DISTRIBUTE_JOBS_PER_THREAD = 0, UNIFORM = 1 };

class JobsQueue {
public:
    void add(const int32_t* lits, size_t n);
    void close() { closed_ = true; }
    bool closed() const { return closed_; }
    size_t n_jobs() const { return off_.size() - 1; }
    size_t size_of_largest_job() const { return largest_; }
    const std::vector<int32_t>& lits() const { return lits_; }
    const std::vector<int32_t>& offsets() const { return off_; }

private:
    std::vector<int32_t> lits_;
    std::vector<int32_t> off_{0};
    size_t largest_ = 0;
    bool closed_ = false;
};

class JobChooser {
public:
    virtual ~JobChooser() = default;
    virtual void evaluate() = 0;
    virtual void getJobs(JobsQueue& queue) = 0;
    virtual size_t get_n_jobs() = 0;
    virtual bool is_evaluated() = 0;
    const std::vector<int32_t>& chosen_vars() const { return chosen_; }

protected:
    // all 2^d sign combinations of the chosen vars, depth first, positive first
    void add_jobs(JobsQueue& queue);
    std::vector<int32_t> chosen_;
};

// MaxClauseJobChooser (src/JobsManager/JobChooser.cu:52-133) with VariableChooser
// (src/JobsManager/VariableChooser.cu:23-40): depth from the thread count, vars ranked
// by the sum of their index over occurrences, ties to the lower var.
class MaxClauseJobChooser : public JobChooser {
public:
    MaxClauseJobChooser(const int32_t* lits, const int64_t* clause_off, int32_t m, int32_t n_vars, int32_t n_dead,
                        int64_t threads, int64_t blocks, ChoosingStrategy strategy);
    void evaluate() override;
    void getJobs(JobsQueue& queue) override { add_jobs(queue); }
    size_t get_n_jobs() override { return (size_t)1 << vars_per_job_; }
    bool is_evaluated() override { return evaluated_; }
    int32_t vars_per_job() const { return vars_per_job_; }

private:
    const int32_t* lits_;
    const int64_t* clause_off_;
    int32_t m_, n_vars_;
    int64_t n_working_vars_, threads_, blocks_;
    ChoosingStrategy strategy_;
    int32_t vars_per_job_ = 0;
    bool evaluated_ = false;
};

// SimpleJobChooser (src/JobsManager/SimpleJobChooser.cu:22-49): the first
// UNIFORM_NUMBER_OF_VARS (7) vars that are not dead.
class SimpleJobChooser : public JobChooser {
public:
    SimpleJobChooser(int32_t n_vars, const std::vector<int32_t>& dead_vars) : n_vars_(n_vars), dead_(dead_vars) {}
    void evaluate() override;
    void getJobs(JobsQueue& queue) override { add_jobs(queue); }
    size_t get_n_jobs() override { return (size_t)1 << chosen_.size(); }
    bool is_evaluated() override { return evaluated_; }

private:
    int32_t n_vars_;
    std::vector<int32_t> dead_;
    bool evaluated_ = false;
};

// Cube split from per-var scores (for example those of the learned branching
// network): the d highest-scoring live vars, ties to the lower var.
class ScoredJobChooser : public JobChooser {
public:
    ScoredJobChooser(const float* scores, int32_t n_vars, const std::vector<int32_t>& dead_vars, int32_t depth)
        : scores_(scores), n_vars_(n_vars), dead_(dead_vars), depth_(depth) {}
    void evaluate() override;
    void getJobs(JobsQueue& queue) override { add_jobs(queue); }
    size_t get_n_jobs() override { return (size_t)1 << chosen_.size(); }
    bool is_evaluated() override { return evaluated_; }

private:
    const float* scores_;
    int32_t n_vars_;
    std::vector<int32_t> dead_;
    int32_t depth_;
    bool evaluated_ = false;
};

// FormulaData::copy_host_clauses_to_dev preprocessing (src/FileManager/FormulaData.cu:82-105).
struct Preprocessed {
    int status = 2;                  // 0 SAT, 1 UNSAT, 2 formula remains
    std::vector<int32_t> lits;
    std::vector<int64_t> clause_off{0};
    std::vector<int32_t> solved;     // literals fixed by unit clauses
};
Preprocessed preprocess(const int32_t* lits, const int64_t* clause_off, int32_t m, int32_t n_vars);

}  // namespace gpsat
